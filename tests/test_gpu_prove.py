"""GPU parity of the prove loop: PoW search, eq / geometric weight kernels,
MLE evaluation, and whole proofs (narg + hints byte for byte) against the
reference's own prover compiled from its sources (oracle/_ref), and against
the committed proof digests made from it.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu
P = 0xFFFFFFFF00000001
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx():
    c = W.Context()
    yield c
    c.close()


def _e3s(rnd, n, base_every=0):
    out = []
    for i in range(n):
        if base_every and i % base_every == 0:
            out.append([rnd.randrange(P), 0, 0])
        else:
            out.append([rnd.randrange(P) for _ in range(3)])
    return np.array(out, dtype=np.uint64).reshape(-1, 3)


# ------------------------------------------------------------------- PoW

def test_pow_golden(ctx, fixtures):
    g = fixtures["pow_indices"]["pow"]
    print("pow golden", flush=True)
    assert ctx.pow_search(bytes([g["fill"]]) * 32, g["threshold"]) == g["nonce"]
    assert ctx.pow_search(bytes(32), 2**64 - 1) == 0


@pytest.mark.parametrize("bits", [1, 4, 8, 12, 16])
def test_pow_vs_oracle(ctx, bits):
    rnd = random.Random(bits)
    for _ in range(3):
        ch = rnd.randbytes(32)
        thr = int(np.ceil(2.0 ** (64 - bits)))
        print("pow", bits, ch.hex(), flush=True)
        assert ctx.pow_search(ch, thr) == oracle.pow_find(ch, thr)


def test_pow_past_first_window(ctx):
    """A threshold that the first 2^22-nonce window rarely meets: the search
    must continue window by window and still return the global minimum."""
    ch = bytes(range(32))
    thr = 2 ** (64 - 23)
    print("pow 23 bits", flush=True)
    n = ctx.pow_search(ch, thr)
    h = oracle.blake3(ch + n.to_bytes(8, "little") + bytes(24))
    assert int.from_bytes(h[:8], "little") <= thr
    # the oracle's serial scan is exact but slow: check minimality below n on a sample
    rnd = random.Random(1)
    for m in [rnd.randrange(n) for _ in range(2000)] if n else []:
        hm = oracle.blake3(ch + m.to_bytes(8, "little") + bytes(24))
        assert int.from_bytes(hm[:8], "little") > thr


# --------------------------------------------------------------- weights

@pytest.mark.parametrize("nv", [0, 1, 3, 4, 9, 10, 11, 16])
def test_eq_accumulate(ctx, nv):
    rnd = random.Random(100 + nv)
    n = 1 << nv
    x = _e3s(rnd, nv)
    s = _e3s(rnd, 1)[0]
    base = _e3s(rnd, n).reshape(-1)
    d = ctx.upload(base)
    print("eq", nv, flush=True)
    ctx.eq_accumulate(d, x, s, accumulate=True)
    got = ctx.download(d, 3 * n)
    assert np.array_equal(got, oracle.eval_eq_ext3(base, x, s))
    ctx.eq_accumulate(d, x, s, accumulate=False)
    got = ctx.download(d, 3 * n)
    assert np.array_equal(got, oracle.eval_eq_ext3(np.zeros(3 * n, dtype=np.uint64), x, s))
    ctx.free(d)


def test_eq_golden(ctx, fixtures):
    for c in fixtures["utils"]["eval_eq"]:
        k = len(c["point"])
        d = ctx.upload(np.zeros(3 << k, dtype=np.uint64))
        pt = np.array([[v, 0, 0] for v in c["point"]], dtype=np.uint64)
        ctx.eq_accumulate(d, pt, [c["scalar"], 0, 0], accumulate=True)
        got = ctx.download(d, 3 << k).reshape(-1, 3)
        assert list(got[:, 0]) == c["result"] and not got[:, 1:].any()
        ctx.free(d)


# (1024, 1500) and (2048, 1100, all base) pass the sweep's 1024 limb products
# per accumulator, so its mid-sweep reductions run
@pytest.mark.parametrize("n,count,base_every", [(1, 1, 2), (2, 3, 2), (8, 3, 2), (1024, 5, 2), (4096, 60, 2),
                                                (1 << 18, 253, 2), (1024, 1500, 2), (2048, 1100, 1)])
def test_geometric_accumulate(ctx, n, count, base_every):
    rnd = random.Random(n * 7 + count)
    pts = _e3s(rnd, count, base_every=base_every)  # every base_every-th point in the base field
    scs = _e3s(rnd, count)
    base = _e3s(rnd, n).reshape(-1)
    d = ctx.upload(base)
    print("geometric", n, count, flush=True)
    ctx.geometric_accumulate(d, n, pts, scs)
    got = ctx.download(d, 3 * n)
    if n * count <= 1 << 16:
        assert np.array_equal(got, oracle.geometric_accumulate_ext3(base, pts, scs))
    else:
        # sampled entries against the closed form sum_j s_j z_j^i
        for i in [0, 1, n // 2 + 3, n - 1]:
            want = base[3 * i:3 * i + 3].copy()
            for z, s in zip(pts, scs):
                want = oracle.e3_add(want, oracle.e3_mul(s, oracle.e3_pow(z, i)))
            assert np.array_equal(got[3 * i:3 * i + 3], want), i
    ctx.free(d)


def test_geometric_golden(ctx, fixtures):
    ga = fixtures["utils"]["geometric_accumulate"]
    lift = lambda v: np.array([[x, 0, 0] for x in v], dtype=np.uint64)
    d = ctx.upload(lift(ga["acc"]).reshape(-1))
    ctx.geometric_accumulate(d, 8, lift(ga["points"]), lift(ga["scalars"]))
    got = ctx.download(d, 24).reshape(-1, 3)
    assert list(got[:, 0]) == ga["result"] and not got[:, 1:].any()
    ctx.free(d)


# nv < 8: one lane per row entry; nv 8-10: rows of 2^L >= 256 (the grid a
# multiple of 2^L / 256); below nv 21 (Ext3) and 22 (base) a lane's loop runs
# single steps only, from there on groups of loads in flight (two at nv 23)
@pytest.mark.parametrize("nv,base", [(0, True), (1, False), (5, True), (7, False), (8, True), (9, False), (10, False),
                                     (12, True), (16, False), (20, True), (21, False), (22, True), (23, True)])
def test_mle_evaluate(ctx, nv, base):
    rnd = random.Random(nv)
    n = 1 << nv
    f, _ = W.lcg_uniform(1000 + nv, n * (1 if base else 3))
    x = _e3s(rnd, nv)
    d = ctx.upload(f)
    print("mle", nv, base, flush=True)
    got = ctx.mle_evaluate(d, base, x)
    assert np.array_equal(got, oracle.mle_evaluate(f, x, base))
    ctx.free(d)


# 1-4 points share a launch (5 and 6 take two); base and Ext3 f, rows of
# 2^L below and above 256, single steps and groups of loads per lane
@pytest.mark.parametrize("nv,base,count", [(0, True, 2), (3, False, 3), (7, True, 4), (9, True, 5), (10, False, 4),
                                           (12, True, 3), (16, False, 2), (20, True, 6), (21, False, 3),
                                           (22, True, 3), (23, True, 4)])
def test_mle_evaluate_many(ctx, nv, base, count):
    rnd = random.Random(100 + nv + count)
    n = 1 << nv
    f, _ = W.lcg_uniform(2000 + nv, n * (1 if base else 3))
    pts = np.array([_e3s(rnd, nv) for _ in range(count)], dtype=np.uint64).reshape(count, nv, 3)
    d = ctx.upload(f)
    print("mle many", nv, base, count, flush=True)
    got = ctx.mle_evaluate_many(d, base, pts)
    for k in range(count):
        assert np.array_equal(got[k], oracle.mle_evaluate(f, pts[k], base)), k
    ctx.free(d)


def test_mle_golden(ctx, fixtures):
    for c in fixtures["utils"]["multilinear_extend"]:
        d = ctx.upload(np.array(c["evals"], dtype=np.uint64))
        pt = np.array([[v, 0, 0] for v in c["point"]], dtype=np.uint64)
        assert list(ctx.mle_evaluate(d, True, pt)) == [c["result"], 0, 0]
        ctx.free(d)


# ------------------------------------------------------------ whole proofs

def _inputs(nv, nforms, batch, seed):
    coeffs, st = W.lcg_uniform(seed, batch << nv)
    pts, _ = W.lcg_uniform(st, 3 * nv * nforms)
    return coeffs, pts.reshape(nforms, nv, 3) if nforms else np.zeros((0, nv, 3), dtype=np.uint64)


def gpu_prove(ctx, nv, coeffs, points, evals, **kw):
    p = W.Params.default(unique_decoding=kw.get("unique", 0), starting_log_inv_rate=kw.get("rate", 1),
                         initial_folding_factor=kw.get("iff", 4), folding_factor=kw.get("ff", 4),
                         security_level=kw.get("sec", 128), pow_bits=kw.get("pow_bits", 20),
                         batch_size=kw.get("batch", 1))
    pid, sid = W.domain_separator_cli(p)
    d = ctx.upload(coeffs)
    with W.Prover(ctx, p, 1 << nv, pid, sid) as pr:
        pr.commit(d, kw.get("batch", 1))
        pr.prove(points, evals)
        out = pr.proof()
    ctx.free(d)
    return out


# (nv, prm) — CLI defaults unless given; small ones also exercise tiny codewords
PROOF_CASES = [
    (8, dict(nforms=1)),
    (10, dict(nforms=1)),
    (12, dict(nforms=2, rate=2)),
    (13, dict(nforms=3, ff=3, iff=3)),
    (14, dict(nforms=0)),
    (12, dict(nforms=0, unique=1)),
    (12, dict(nforms=2, unique=1)),
    (11, dict(nforms=2, batch=3)),
    (16, dict(nforms=3)),
    (16, dict(nforms=1, ff=2, iff=5, sec=100, pow_bits=16)),
    (18, dict(nforms=3)),
    (20, dict(nforms=3)),
]


def _case_id(c):
    nv, kw = c
    return f"nv{nv}-" + "-".join(f"{k}{v}" for k, v in sorted(kw.items()))


@pytest.mark.skipif("not __import__('oracle').ref()")
@pytest.mark.parametrize("case", PROOF_CASES, ids=_case_id)
def test_proof_vs_reference(ctx, case):
    nv, kw = case
    kw = dict(kw)
    nforms = kw.pop("nforms")
    coeffs, pts = _inputs(nv, nforms, kw.get("batch", 1), 0x5EED0000 + nv)
    print("proof", nv, nforms, kw, flush=True)
    ev, narg, hints = oracle.ref_prove(nv, coeffs, pts, **kw)
    # the constraint values as the GPU computes them
    d = ctx.upload(coeffs)
    n = 1 << nv
    for f in range(nforms):
        for v in range(kw.get("batch", 1)):
            got = ctx.mle_evaluate(d + 8 * n * v, True, pts[f])
            assert np.array_equal(got, ev[f * kw.get("batch", 1) + v])
    ctx.free(d)
    g_narg, g_hints = gpu_prove(ctx, nv, coeffs, pts, ev, **kw)
    assert len(g_narg) == len(narg) and len(g_hints) == len(hints)
    assert g_narg == narg
    assert g_hints == hints


def test_proof_digests(ctx):
    """Committed digests of reference proofs (tests/golden/proofs.json, made by
    tests/golden/make_proof_fixtures.py from the reference build)."""
    path = os.path.join(HERE, "golden", "proofs.json")
    with open(path) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        nv, kw = c["nv"], dict(c["params"])
        nforms = kw.pop("nforms")
        coeffs, pts = _inputs(nv, nforms, kw.get("batch", 1), c["seed"])
        ev = np.array(c["evaluations"], dtype=np.uint64).reshape(-1, 3)
        print("digest", nv, c["params"], flush=True)
        narg, hints = gpu_prove(ctx, nv, coeffs, pts, ev, **kw)
        assert hashlib.sha256(narg).hexdigest() == c["narg_sha256"]
        assert hashlib.sha256(hints).hexdigest() == c["hints_sha256"]
        assert (len(narg), len(hints)) == (c["narg_len"], c["hints_len"])


# Proofs the reference verifier must accept, at sizes and parameters beyond the
# byte-for-byte comparison above (nv 24 included); the claimed values come from
# the GPU's own MLE evaluation.
VERIFY_CASES = [
    (12, dict(nforms=1)),
    (16, dict(nforms=2, rate=2)),
    (20, dict(nforms=3)),
    (22, dict(nforms=3)),
    (24, dict(nforms=1)),
    (14, dict(nforms=2, batch=4)),
    (15, dict(nforms=1, unique=1)),
    (17, dict(nforms=2, ff=3, iff=5)),
    (18, dict(nforms=2, hash_id=2)),
]


@pytest.mark.skipif("not __import__('oracle').ref()")
@pytest.mark.parametrize("case", VERIFY_CASES, ids=_case_id)
def test_gpu_proof_verifies(ctx, case):
    nv, kw = case
    kw = dict(kw)
    nforms = kw.pop("nforms")
    hid = kw.pop("hash_id", 1)
    batch = kw.get("batch", 1)
    coeffs, pts = _inputs(nv, nforms, batch, 0x7E51F000 + nv)
    print("verify", nv, nforms, kw, hid, flush=True)
    n = 1 << nv
    d = ctx.upload(coeffs)
    ev = np.array([ctx.mle_evaluate(d + 8 * n * v, True, pts[f]) for f in range(nforms) for v in range(batch)],
                  dtype=np.uint64).reshape(-1, 3)
    p = W.Params.default(unique_decoding=kw.get("unique", 0), starting_log_inv_rate=kw.get("rate", 1),
                         initial_folding_factor=kw.get("iff", 4), folding_factor=kw.get("ff", 4),
                         batch_size=batch, hash_id=hid)
    pid, sid = W.domain_separator_cli(p)
    with W.Prover(ctx, p, n, pid, sid) as pr:
        pr.commit(d, batch)
        pr.prove(pts, ev)
        narg, hints = pr.proof()
    ctx.free(d)
    assert oracle.ref_verify(nv, pts, ev, narg, hints, hash_id=hid, **kw)
    b = bytearray(narg)
    b[len(b) // 2] ^= 0x10
    assert not oracle.ref_verify(nv, pts, ev, bytes(b), hints, hash_id=hid, **kw)
