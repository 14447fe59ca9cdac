"""Batched commits in one graph launch, and the LDE autotuner: every radix
split gives the same codeword, and every graph replay the same roots as the
eager run and as per-polynomial commits of the reference build / oracle."""
import random

import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = W.Context()
    yield c
    c.close()


def _commit_ref(coeffs, rate=2, ff=4):
    """Root of one polynomial's commit: the reference build when present,
    else the oracle's restatement."""
    n = coeffs.size
    depth = 1 << ff
    cw = (n // depth) << rate
    R = oracle.ref()
    if R is not None:
        root = np.zeros(32, dtype=np.uint8)
        R.ref_commit_base(oracle._ptr64(coeffs), n, cw, depth, None, oracle._ptr8(root))
        return root.tobytes()
    m = oracle.rs_encode(coeffs, 1, n, cw, depth)
    leaves = oracle.commit_leaves_base(m, cw, depth)
    return oracle.merkle_build(leaves, cw)[-32:].tobytes()


def _polys(sizes, seed):
    out, st = [], seed
    for s in sizes:
        c, st = W.lcg_uniform(st, s)
        out.append(c)
    return out


def test_batch_roots_and_replays(ctx):
    rnd = random.Random(3)
    sizes = [1 << rnd.randint(8, 16) for _ in range(20)] + [1 << 12, 1 << 12]
    polys = _polys(sizes, 0xBA7C)
    print("batch", sizes, flush=True)
    with W.Batch(ctx, sizes, log_inv_rate=2, folding_factor=4, branches=4) as b:
        for k, c in enumerate(polys):
            b.upload(k, c)
        b.run(graph=False)
        eager = b.roots()
        assert eager == [_commit_ref(c) for c in polys]
        # the codeword of one polynomial against the oracle
        k = sizes.index(1 << 12)
        assert np.array_equal(b.matrix(k), oracle.rs_encode(polys[k], 1, sizes[k], (sizes[k] // 16) << 2, 16))
        # whole node arrays (the trees are built level by level across the
        # batch): one tree per step shape of the build, against the oracle
        for nv in sorted({s.bit_length() - 1 for s in sizes}):
            k = sizes.index(1 << nv)
            rows = sizes[k] // 4
            want = oracle.merkle_build(oracle.commit_leaves_base(b.matrix(k), rows, 16), rows)
            assert np.array_equal(b.nodes(k), want), nv
        # graph replays, with the program's own eager work in between
        scratch = ctx.upload(polys[0])
        d_m = ctx.alloc(8 * (sizes[0] // 4) * 16)
        d_n = ctx.alloc(32 * W.merkle_num_nodes(sizes[0] // 4))
        for rep in range(3):
            b.run(graph=True)
            assert b.roots() == eager, rep
            ctx.commit(scratch, 1, sizes[0], sizes[0] // 4, 16, d_m, d_n)
            ctx.sync()
        # new coefficients are picked up by the next replay
        newc, _ = W.lcg_uniform(77, sizes[1])
        b.upload(1, newc)
        b.run(graph=True)
        r = b.roots()
        assert r[1] == _commit_ref(newc) and r[0] == eager[0] and r[2:] == eager[2:]


@pytest.mark.parametrize("rate,lgs", [(2, list(range(22, 13, -1)) + [14, 12]), (1, [18, 15, 14, 11]),
                                       (3, [21, 17, 14]), (4, [22, 16, 14])])
def test_batch_level_synchronous(ctx, rate, lgs):
    """The level-synchronous form (one first-pass launch and one second-pass
    launch over all polynomials, trees level by level): every root equals the
    reference build's commit of that polynomial alone, on the eager run and on
    graph replays, at every first-pass size (2^11 .. 2^18 coefficients per
    column) and one-pass size (2^10), beside a polynomial too small for it
    (committed per polynomial in the same graph); whole node arrays and
    codewords against the oracle for a few."""
    sizes = [1 << lg for lg in lgs]
    polys = _polys(sizes, 0x1E7E1 + rate)  # rate 4: 2^22 coefficients -> 2^22 rows, 16 cosets
    print("level batch rate", rate, sizes, flush=True)
    want = [_commit_ref(c, rate=rate) for c in polys]
    with W.Batch(ctx, sizes, log_inv_rate=rate, folding_factor=4, branches=2, autotune=False, form=1) as b:
        assert b.form()[0] == 1
        for k, c in enumerate(polys):
            b.upload(k, c)
        b.run(graph=False)
        assert b.roots() == want
        for rep in range(2):
            b.run(graph=True)
            assert b.roots() == want, rep
        for k in sorted(range(len(sizes)), key=lambda k: sizes[k])[:3]:
            rows = (sizes[k] // 16) << rate
            m = oracle.rs_encode(polys[k], 1, sizes[k], rows, 16)
            assert np.array_equal(b.matrix(k), m), sizes[k]
            assert np.array_equal(b.nodes(k), oracle.merkle_build(oracle.commit_leaves_base(m, rows, 16), rows))


def _force(rec_key, lg):
    W.tuning_import(np.array(list(rec_key) + [lg], dtype=np.uint64))


def test_batch_sha256(ctx):
    """A batch made on a SHA-256 context commits with SHA-256 leaves and trees,
    eagerly and on graph replays, also after the context goes back to BLAKE3."""
    sizes = [1 << 10, 1 << 14, 1 << 12]
    polys = _polys(sizes, 0x5A2B)
    print("batch sha256", sizes, flush=True)
    ctx.set_hash(W.HASH_SHA2)
    try:
        b = W.Batch(ctx, sizes, log_inv_rate=2, folding_factor=4, branches=2)
    finally:
        ctx.set_hash(W.HASH_BLAKE3)
    with b:
        assert b.form()[0] == 0  # the level-synchronous form is BLAKE3 only
        for k, c in enumerate(polys):
            b.upload(k, c)
        want = []
        for k, c in enumerate(polys):
            cw = (sizes[k] // 16) << 2
            m = oracle.rs_encode(c, 1, sizes[k], cw, 16)
            want.append(oracle.sha256_merkle_build(oracle.sha256_leaves(m, cw, 16), cw)[-32:].tobytes())
        b.run(graph=False)
        assert b.roots() == want
        for rep in range(2):
            b.run(graph=True)
            assert b.roots() == want, rep


@pytest.mark.parametrize("m,n,ext3", [(1 << 14, 1 << 16, False), (1 << 12, 1 << 17, True), (1 << 16, 1 << 18, False)])
def test_every_split_same_codeword(ctx, m, n, ext3):
    depth = 16
    coeffs, _ = W.lcg_uniform(m + n, m * depth * (3 if ext3 else 1))
    d_in = ctx.upload(coeffs)
    d_out = ctx.alloc(8 * n * depth * (3 if ext3 else 1))
    ctx.rs_encode(d_in, 1, m * depth, n, depth, d_out, ext3=ext3)
    want = ctx.download(d_out, n * depth * (3 if ext3 else 1))
    if not ext3:
        assert np.array_equal(want, oracle.rs_encode(coeffs, 1, m * depth, n, depth))
    else:
        assert np.array_equal(want, oracle.rs_encode_ext3(coeffs, 1, m * depth, n, depth))
    key = (m, n, depth * 3, 3, 1, 1, 1, 0) if ext3 else (m, n, depth, 1, 1, 0, 1, 0)
    default = min(10, m.bit_length() - 1)
    tried = 0
    for lg in range(default, 0, -1):
        if (m >> lg) > 1024:
            break
        _force(key, lg)
        ctx.rs_encode(d_in, 1, m * depth, n, depth, d_out, ext3=ext3)
        got = ctx.download(d_out, n * depth * (3 if ext3 else 1))
        assert np.array_equal(got, want), lg
        tried += 1
    assert tried >= 2
    # both forms of the fused-leaves second pass (bits 8-15 of the record)
    for form in (1, 2):
        _force(key, default | (form << 8))
        ctx.rs_encode(d_in, 1, m * depth, n, depth, d_out, ext3=ext3)
        assert np.array_equal(ctx.download(d_out, n * depth * (3 if ext3 else 1)), want), form
    _force(key, default)
    ms, lg = W.autotune_lde(ctx, m, n, depth, ext3)
    assert ms > 0 and (m >> lg) <= 1024
    ctx.rs_encode(d_in, 1, m * depth, n, depth, d_out, ext3=ext3)
    assert np.array_equal(ctx.download(d_out, n * depth * (3 if ext3 else 1)), want)
    ctx.free(d_in)
    ctx.free(d_out)


BATCH_SEED = 0x5748495200000003  # the batched-commit workload's seed (bench.py uses the same)


def test_batch_workload_64_polys(ctx):
    """The batched-commit workload itself: 64 polynomials with num_variables
    uniform on [14, 22] from the seed (about 2^25.6 coefficients), rate 1/4,
    folding factor 4, every shape autotuned.  All 64 roots equal the
    reference build's commit of each polynomial (irs_commit.hpp:344-460), on
    the eager run and on three graph replays; the largest polynomial's
    codeword equals the reference's interleaved_rs_encode, and a 2^16 one's
    the oracle's."""
    rnd = random.Random(BATCH_SEED)
    sizes = [1 << rnd.randint(14, 22) for _ in range(64)]
    polys = _polys(sizes, BATCH_SEED)
    print("batch workload", sum(sizes), "coefficients", flush=True)
    want = [_commit_ref(c) for c in polys]
    # both batch forms, each eager and on graph replays
    for form in (0, 1):
        with W.Batch(ctx, sizes, log_inv_rate=2, folding_factor=4, branches=4, autotune=False, form=form) as b:
            assert b.form()[0] == form
            for k, c in enumerate(polys):
                b.upload(k, c)
            b.run(graph=False)
            assert b.roots() == want, form
            for rep in range(2):
                b.run(graph=True)
                assert b.roots() == want, (form, rep)
    with W.Batch(ctx, sizes, log_inv_rate=2, folding_factor=4, branches=4, autotune=True) as b:
        f, per_ms, level_ms = b.form()
        print("batch forms: per-polynomial", per_ms, "ms, level-synchronous", level_ms, "ms, chose", f, flush=True)
        assert per_ms > 0 and level_ms > 0 and f == (1 if level_ms <= per_ms else 0)
        for k, c in enumerate(polys):
            b.upload(k, c)
        b.run(graph=False)
        assert b.roots() == want
        for rep in range(3):
            b.run(graph=True)
            assert b.roots() == want, rep
        big = max(range(64), key=lambda k: sizes[k])
        n, cw = sizes[big], (sizes[big] // 16) << 2
        R = oracle.ref()
        if R is not None:
            ref_m = np.zeros(cw * 16, dtype=np.uint64)
            R.ref_rs_encode(oracle._ptr64(polys[big]), 1, n, cw, 16, oracle._ptr64(ref_m))
        else:
            ref_m = oracle.rs_encode(polys[big], 1, n, cw, 16)
        assert np.array_equal(b.matrix(big), ref_m)
        mid = next((k for k in range(64) if sizes[k] == 1 << 16), None)
        if mid is not None:
            cw = (1 << 16) // 16 << 2
            assert np.array_equal(b.matrix(mid), oracle.rs_encode(polys[mid], 1, 1 << 16, cw, 16))
