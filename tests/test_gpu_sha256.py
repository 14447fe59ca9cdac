"""The SHA-256 hash engine on the GPU (hash_id 2, the CLI's --hash Sha2):
hash_many, Merkle trees and open paths, row leaves of a commit, the
proof-of-work search and whole proofs, each against the oracle (hashlib's
SHA-256, pinned on the reference's sha2 golden vectors) or the reference's
own prover built from its sources.

Reference behaviour: hash/sha2_engine.hpp:47-81 (SHA-256 of each message),
cuda_integration.hpp:888-925 / 995-1045 (gpu_sha256_merkle_tree /
_open_path), tests/test_cuda_merkle_sha256.cpp (tree sizes 0, 1, 2, 3, 17,
1024, 65536).
"""
import random

import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu
SHA2 = W.HASH_SHA2


def rand(seed, n):
    return np.random.default_rng(seed).integers(0, oracle.P, size=n, dtype=np.uint64)


@pytest.fixture
def ctx():
    c = W.Context()
    c.set_hash(SHA2)
    yield c
    c.close()


@pytest.mark.parametrize("size", [0, 1, 31, 55, 56, 63, 64, 65, 119, 120, 128, 1000, 1025])
def test_sha256_hash_many(size):
    count = 37
    data = bytes(np.random.default_rng(size).integers(0, 256, size * count, dtype=np.uint8))
    print("sha256 hash_many", size, flush=True)
    got = W.sha256_hash_many(size, data, count)
    for i in range(count):
        assert got[32 * i:32 * i + 32].tobytes() == oracle.sha256(data[i * size:(i + 1) * size])


def test_sha256_golden(fixtures):
    n = 0
    for c in fixtures["hash"]:
        if c["engine"] != "sha2":
            continue
        got = W.sha256_hash_many(c["size"], bytes.fromhex(c["input"]), c["count"])
        assert [got[32 * i:32 * i + 32].tobytes().hex() for i in range(c["count"])] == c["digests"]
        n += 1
    assert n


def test_sha256_merkle_golden(fixtures):
    for c in fixtures["merkle_tree"]:
        if c["hash"] != "sha2":
            continue
        leaves = np.frombuffer(bytes.fromhex(c["leaves"]), dtype=np.uint8)
        root, hints = W.merkle_open_path(leaves, c["num_leaves"], c["indices"], hash=SHA2)
        assert root.tobytes().hex() == c["root"]
        assert [hints[32 * i:32 * i + 32].tobytes().hex() for i in range(len(c["hints"]))] == c["hints"]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 1024, 4097, 65536, (1 << 18) + 5])
def test_sha256_merkle_tree(n):
    leaves = np.random.default_rng(n).integers(0, 256, 32 * n, dtype=np.uint8)
    print("sha256 tree", n, flush=True)
    root, nodes = W.merkle_tree(leaves, n, hash=SHA2)
    want = oracle.sha256_merkle_build(leaves, n)
    assert np.array_equal(nodes, want)
    assert root.tobytes() == want[-32:].tobytes()
    root_only, _ = W.merkle_tree(leaves, n, full=False, hash=SHA2)
    assert root_only.tobytes() == root.tobytes()


@pytest.mark.parametrize("n", [1, 2, 5, 1000, 1 << 16])
def test_sha256_merkle_open_path(n):
    rng = np.random.default_rng(n + 1)
    leaves = rng.integers(0, 256, 32 * n, dtype=np.uint8)
    idx = list(rng.integers(0, n, 32)) + [0, n - 1, n - 1, int(rng.integers(0, n))]
    print("sha256 open", n, flush=True)
    root, hints = W.merkle_open_path(leaves, n, idx, hash=SHA2)
    nodes = oracle.sha256_merkle_build(leaves, n)
    assert root.tobytes() == nodes[-32:].tobytes()
    assert np.array_equal(hints, oracle.merkle_open(nodes, n, idx))


def test_sha256_leaves_row_lengths(ctx):
    """Row leaves of 1..200 limbs: every padding case of the last block
    (0-7 limbs left over, the length in a second block after 7)."""
    for rows, row_u64 in ((1000, 16), (77, 5), (33, 48), (9, 200), (4, 1), (50, 7), (50, 15), (20, 8)):
        print("sha256 leaves", rows, row_u64, flush=True)
        m = rand(rows * row_u64, rows * row_u64)
        dm = ctx.upload(m)
        dl = ctx.alloc(32 * rows)
        ctx.commit_leaves(dm, rows, row_u64, dl)
        got = ctx.download(dl, 32 * rows, np.uint8)
        assert np.array_equal(got, oracle.sha256_leaves(m, rows, row_u64))


@pytest.mark.parametrize("shape", [(1, 1 << 12, 1 << 14, 16), (2, 1 << 10, 1 << 10, 4)])
def test_sha256_dev_commit(ctx, shape):
    num_polys, poly_size, cw, depth = shape
    x = rand(5 + sum(shape), num_polys * poly_size)
    print("sha256 commit", shape, flush=True)
    dx = ctx.upload(x)
    cols = num_polys * depth
    dm = ctx.alloc(8 * cw * cols)
    dn = ctx.alloc(32 * W.merkle_num_nodes(cw))
    root = ctx.commit(dx, num_polys, poly_size, cw, depth, dm, dn)
    m = ctx.download(dm, cw * cols)
    nodes = ctx.download(dn, 32 * W.merkle_num_nodes(cw), np.uint8)
    want = oracle.rs_encode(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(m, want)
    wn = oracle.sha256_merkle_build(oracle.sha256_leaves(want, cw, cols), cw)
    assert np.array_equal(nodes, wn)
    assert root == wn[-32:].tobytes()
    # the same context back on BLAKE3 gives the BLAKE3 commitment
    ctx.set_hash(W.HASH_BLAKE3)
    root_b = ctx.commit(dx, num_polys, poly_size, cw, depth, dm, dn)
    assert root_b == oracle.merkle_build(oracle.commit_leaves_base(want, cw, cols), cw)[-32:].tobytes()


@pytest.mark.parametrize("poly_size,cw", [(1 << 12, 1 << 11), (1 << 15, 1 << 13), (1 << 16, 1 << 16)])
def test_sha256_dev_commit_ext3(ctx, poly_size, cw):
    """Ext3 round commits with SHA-256 trees: from 2048 coefficients per column
    the 384-byte row leaves are hashed inside the encode (six data blocks over
    the three 16-column sub-tiles, then the padding block); every node of the
    tree against the oracle."""
    depth = 16
    x = rand(17 + poly_size + cw, 3 * poly_size)
    print("sha256 ext3 commit", poly_size, cw, flush=True)
    dx = ctx.upload(x)
    dm = ctx.alloc(24 * cw * depth)
    dn = ctx.alloc(32 * W.merkle_num_nodes(cw))
    root = ctx.commit(dx, 1, poly_size, cw, depth, dm, dn, ext3=True)
    m = ctx.download(dm, 3 * cw * depth)
    nodes = ctx.download(dn, 32 * W.merkle_num_nodes(cw), np.uint8)
    want = oracle.rs_encode_ext3(x, 1, poly_size, cw, depth)
    assert np.array_equal(m, want)
    wn = oracle.sha256_merkle_build(oracle.sha256_leaves(want, cw, 3 * depth), cw)
    assert np.array_equal(nodes, wn)
    assert root == wn[-32:].tobytes()


def test_sha256_commit_golden_whir(ctx, fixtures):
    """The reference's golden_whir proofs (dump_whir.cpp: SHA-256 trees, depth
    1, codeword 2x the vector size): the device commit of their input vectors
    gives their witness matrix, and its root opens their narg string."""
    for c in fixtures["whir"]:
        nv, vs = c["num_vectors"], c["vector_size"]
        print("sha256 golden_whir commit", nv, vs, flush=True)
        dx = ctx.upload(np.array(sum(c["vectors"], []), dtype=np.uint64))
        dm = ctx.alloc(8 * 2 * vs * nv)
        dn = ctx.alloc(32 * W.merkle_num_nodes(2 * vs))
        root = ctx.commit(dx, nv, vs, 2 * vs, 1, dm, dn)
        assert list(ctx.download(dm, 2 * vs * nv)) == c["witness_matrix"]
        assert root.hex() == c["proof_narg"][:64]


@pytest.mark.parametrize("bits", [1, 4, 8, 12])
def test_sha256_pow(ctx, bits):
    rnd = random.Random(100 + bits)
    for _ in range(3):
        ch = rnd.randbytes(32)
        thr = int(np.ceil(2.0 ** (64 - bits)))
        print("sha256 pow", bits, ch.hex(), flush=True)
        assert ctx.pow_search(ch, thr) == oracle.sha256_pow_find(ch, thr)


def test_set_hash_rejects_unknown(ctx):
    with pytest.raises(W.WhirGpuError):
        ctx.set_hash(3)


# ------------------------------------------------------------ whole proofs

SHA2_PROOF_CASES = [
    (10, dict(nforms=1)),
    (12, dict(nforms=2, rate=2)),
    (11, dict(nforms=2, batch=3)),
    (13, dict(nforms=3, ff=3, iff=3)),
    (16, dict(nforms=3)),
]


@pytest.mark.skipif("not __import__('oracle').ref()")
@pytest.mark.parametrize("case", SHA2_PROOF_CASES, ids=lambda c: f"nv{c[0]}-" + "-".join(f"{k}{v}" for k, v in
                                                                                         sorted(c[1].items())))
def test_sha256_proof_vs_reference(case):
    nv, kw = case
    kw = dict(kw)
    nforms = kw.pop("nforms")
    batch = kw.get("batch", 1)
    coeffs, st = W.lcg_uniform(0x5A2A0000 + nv, batch << nv)
    pts, _ = W.lcg_uniform(st, 3 * nv * nforms)
    pts = pts.reshape(nforms, nv, 3)
    print("sha256 proof", nv, nforms, kw, flush=True)
    ev, narg, hints = oracle.ref_prove(nv, coeffs, pts, hash_id=2, **kw)
    p = W.Params.default(starting_log_inv_rate=kw.get("rate", 1), initial_folding_factor=kw.get("iff", 4),
                         folding_factor=kw.get("ff", 4), batch_size=batch, hash_id=2)
    pid, sid = W.domain_separator_cli(p)
    with W.Context() as c:
        d = c.upload(coeffs)
        with W.Prover(c, p, 1 << nv, pid, sid) as pr:
            pr.commit(d, batch)
            pr.prove(pts, ev)
            g_narg, g_hints = pr.proof()
    assert g_narg == narg
    assert g_hints == hints


@pytest.mark.skipif("not __import__('oracle').ref()")
def test_sha256_served_vs_reference():
    params = W.Params.default(hash_id=2)
    with W.Server(params, num_slots=2, hbm_budget=1 << 30) as srv:
        cases = []
        for i, nv in enumerate([12, 14]):
            c, st = W.lcg_uniform(0x5E2A00 + i, 1 << nv)
            p, _ = W.lcg_uniform(st, 3 * nv * 2)
            p = p.reshape(2, nv, 3)
            print("sha256 served", nv, flush=True)
            cases.append((nv, c, p, srv.submit(1 << nv, c, p)))
        for nv, c, p, rid in cases:
            _, narg, hints = oracle.ref_prove(nv, c, p, hash_id=2)
            assert srv.wait(rid)[:2] == (narg, hints)
