"""GPU parity: every HIP path against the CPU oracle, bit-exact.

Case design follows the reference's own CPU-vs-GPU tests
(cpp/tests/test_cuda_ntt.cpp:75-107, test_cuda_rs_encode.cpp:170-219,
test_cuda_merkle_sha256.cpp:60-87, test_cuda_merkle_path.cpp:56-124):
NTT sizes 2..2^20 x batch 1 and 3 with special values, RS shapes
(poly, cw, depth, polys), trees of 0/1/2/3/17/1024/65536 leaves, openings
with duplicate and unsorted indices; plus the golden fixtures.
"""
import numpy as np
import pytest

import whirplonk_amd as W

pytestmark = pytest.mark.gpu

P = 0xFFFFFFFF00000001
SPECIAL = [0, 1, P - 1, P - 2, 2**32 - 1, 2**32, 2**64 - 1 - 2**32, 7]  # test_cuda_ntt.cpp:89-107 (canonical)


@pytest.fixture(scope="module", autouse=True)
def device():
    assert W.device_available(), "no gfx950 device visible: the GPU tests need an MI355X"
    W.load().whir_set_gpu_ntt_threshold(0)
    yield
    W.load().whir_set_gpu_ntt_threshold(65536)


def rand(seed, n):
    from oracle import lcg

    return lcg(seed, n, 1)[0]


# ------------------------------------------------------------------- NTT

@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 5, 6, 8, 10, 11, 12, 16, 20])
@pytest.mark.parametrize("batch", [1, 3])
def test_ntt_batch(orc, log_n, batch):
    n = 1 << log_n
    x = rand(1000 + log_n * 7 + batch, n * batch)
    x[: min(len(SPECIAL), x.size)] = np.array(SPECIAL, dtype=np.uint64)[: min(len(SPECIAL), x.size)]
    got = W.ntt_batch(x, n)
    assert np.array_equal(got, orc.ntt_batch(x, n))
    inv = W.ntt_batch(got, n, inverse=True)
    assert np.array_equal(inv, orc.intt_batch(got, n))


@pytest.mark.parametrize("log_n,batch", [(21, 3), (22, 1), (24, 1)])
def test_ntt_long(orc, log_n, batch):
    # longer than 2^20: E-point DFTs with the twist, then the 2^20-point NTTs
    n = 1 << log_n
    x = rand(2000 + log_n, n * batch)
    x[: len(SPECIAL)] = np.array(SPECIAL, dtype=np.uint64)
    got = W.ntt_batch(x, n)
    assert np.array_equal(got, orc.ntt_batch(x, n))
    assert np.array_equal(W.ntt_batch(got, n, inverse=True), orc.intt_batch(got, n))


@pytest.mark.parametrize("log_n", [25, 26])
def test_ntt_long_round_trip(log_n):
    # at the largest sizes, properties that fix the transform: the unscaled
    # inverse of the forward transform is n * x (checked on a sample of
    # entries), and the transform of the unit vector e_1 is the powers of a
    # primitive n-th root of unity
    n = 1 << log_n
    x = rand(3000 + log_n, n)
    back = W.ntt_batch(W.ntt_batch(x, n), n, inverse=True)
    idx = np.random.default_rng(log_n).integers(0, n, 20000)
    assert all(int(back[i]) == int(x[i]) * n % P for i in idx)
    e1 = np.zeros(n, dtype=np.uint64)
    e1[1] = 1
    y = W.ntt_batch(e1, n)
    w = int(y[1])
    assert int(y[0]) == 1 and pow(w, n, P) == 1 and pow(w, n // 2, P) == P - 1
    assert all(int(y[i]) == pow(w, int(i), P) for i in list(idx[:2000]) + [n - 1])


def test_ntt_special_values(orc):
    n = 1 << 12
    x = np.array([SPECIAL[i % len(SPECIAL)] for i in range(n)], dtype=np.uint64)
    assert np.array_equal(W.ntt_batch(x, n), orc.ntt_batch(x, n))
    for v in SPECIAL:
        x = np.full(1024, v, dtype=np.uint64)
        assert np.array_equal(W.ntt_batch(x, 1024), orc.ntt_batch(x, 1024))


# ------------------------------------------------------------- RS encode

RS_CASES = [
    # (num_polys, poly_size, codeword_length, depth)
    (1, 64, 256, 1), (2, 256, 1024, 4), (4, 1024, 4096, 4), (4, 4096, 16384, 4),   # test_cuda_rs_encode.cpp:170-219
    (1, 2, 4, 1), (1, 8, 16, 1), (3, 4, 8, 1), (4, 8, 16, 1),                      # golden_whir shapes
    (1, 1 << 16, 1 << 14, 16),                                                      # nv=16, rate 1/4, ff=4
    (1, 16, 2048, 16), (1, 1 << 12, 1 << 12, 16), (5, 512, 256, 4), (2, 1 << 14, 1 << 16, 8),
    (1, 1 << 20, 1 << 18, 4),
    (1, 1 << 10, 1 << 12, 16), (2, 1 << 9, 1 << 14, 16), (1, 1 << 8, 1 << 15, 8),     # one pass, m = 64, 32, 32
]


@pytest.mark.parametrize("shape", RS_CASES)
def test_rs_encode(orc, shape):
    num_polys, poly_size, cw, depth = shape
    x = rand(sum(shape), num_polys * poly_size)
    want = orc.rs_encode(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(W.interleaved_rs_encode(x, num_polys, poly_size, cw, depth), want)
    m, leaves = W.rs_encode_leaves(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(m, want)
    assert np.array_equal(leaves, orc.commit_leaves_base(want, cw, num_polys * depth))


@pytest.mark.parametrize("shape", [(1, 64, 256, 4), (2, 256, 1024, 16), (1, 16, 2048, 16), (3, 1024, 512, 4),
                                   # the last rounds' shapes: m = 4, 16, 32, 64 per column, 48 base columns
                                   (1, 64, 1 << 12, 16), (1, 256, 1 << 14, 16), (1, 512, 1 << 12, 16),
                                   (1, 1024, 1 << 13, 16),
                                   # 24 MiB AoS codeword: split into its components out of two staging chunks
                                   (1, 1 << 14, 1 << 16, 16)])
def test_rs_encode_ext3(orc, shape):
    num_polys, poly_size, cw, depth = shape
    x = rand(77 + sum(shape), 3 * num_polys * poly_size)
    want = orc.rs_encode_ext3(x, num_polys, poly_size, cw, depth)
    m, leaves = W.rs_encode_leaves_ext3(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(m, want)
    assert np.array_equal(leaves, orc.commit_leaves_ext3(want, cw, num_polys * depth))


def test_rs_encode_golden_whir(fixtures):
    for c in fixtures["whir"]:
        nv, vs = c["num_vectors"], c["vector_size"]
        coeffs = np.array(sum(c["vectors"], []), dtype=np.uint64)
        assert list(W.interleaved_rs_encode(coeffs, nv, vs, 2 * vs, 1)) == c["witness_matrix"]


def test_rs_rejects_bad_shapes():
    x = np.zeros(12, dtype=np.uint64)
    with pytest.raises(W.WhirGpuError) as e:
        W.interleaved_rs_encode(x, 1, 12, 16, 1)   # message length not a power of two
    assert e.value.status > 0
    with pytest.raises(W.WhirGpuError):
        W.interleaved_rs_encode(x, 1, 12, 16, 5)   # poly_size not a multiple of depth


# ------------------------------------------------------------------ hash

@pytest.mark.parametrize("size", [0, 1, 31, 63, 64, 65, 128, 384, 1000, 1024, 1025, 2048, 3000, 5000])
def test_blake3_hash_many(orc, size):
    count = 37
    data = bytes(np.random.default_rng(size).integers(0, 256, size * count, dtype=np.uint8))
    got = W.blake3_hash_many(size, data, count)
    for i in range(count):
        assert got[32 * i:32 * i + 32].tobytes() == orc.blake3(data[i * size:(i + 1) * size])


def test_blake3_golden(fixtures):
    for c in fixtures["hash"]:
        if c["engine"] != "blake3":
            continue
        got = W.blake3_hash_many(c["size"], bytes.fromhex(c["input"]), c["count"])
        assert [got[32 * i:32 * i + 32].tobytes().hex() for i in range(c["count"])] == c["digests"]


def test_leaves_row_lengths(orc):
    with W.Context() as ctx:
        for rows, row_u64 in ((1000, 16), (77, 5), (33, 48), (9, 200), (4, 1)):
            m = rand(rows * row_u64, rows * row_u64)
            dm = ctx.upload(m)
            dl = ctx.alloc(32 * rows)
            ctx.commit_leaves(dm, rows, row_u64, dl)
            got = ctx.download(dl, 32 * rows, np.uint8)
            assert np.array_equal(got, orc.commit_leaves_base(m, rows, row_u64))


# ---------------------------------------------------------------- Merkle

@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 1024, 4097, 65536, (1 << 18) + 5])
def test_merkle_tree(orc, n):
    leaves = np.random.default_rng(n).integers(0, 256, 32 * n, dtype=np.uint8)
    root, nodes = W.merkle_tree(leaves, n)
    want = orc.merkle_build(leaves, n)
    assert np.array_equal(nodes, want)
    assert root.tobytes() == want[-32:].tobytes()


def test_merkle_golden(fixtures):
    for c in fixtures["merkle_tree"]:
        if c["hash"] != "blake3":
            continue
        leaves = np.frombuffer(bytes.fromhex(c["leaves"]), dtype=np.uint8)
        root, hints = W.merkle_open_path(leaves, c["num_leaves"], c["indices"])
        assert root.tobytes().hex() == c["root"]
        assert [hints[32 * i:32 * i + 32].tobytes().hex() for i in range(len(c["hints"]))] == c["hints"]


@pytest.mark.parametrize("n", [1, 2, 5, 1000, 1 << 18])
def test_merkle_open_path(orc, n):
    rng = np.random.default_rng(n + 1)
    leaves = rng.integers(0, 256, 32 * n, dtype=np.uint8)
    idx = list(rng.integers(0, n, 32)) + [0, n - 1, n - 1, int(rng.integers(0, n))]  # duplicates, unsorted
    root, hints = W.merkle_open_path(leaves, n, idx)
    nodes = orc.merkle_build(leaves, n)
    assert root.tobytes() == nodes[-32:].tobytes()
    assert np.array_equal(hints, orc.merkle_open(nodes, n, idx))
    lh = np.concatenate([leaves[32 * i:32 * i + 32] for i in idx])
    assert orc.merkle_verify(root.tobytes(), n, idx, lh, hints)


# -------------------------------------------------------- commit (device)

@pytest.mark.parametrize("shape", [(1, 1 << 16, 1 << 14, 16), (1, 1 << 12, 1 << 14, 16), (2, 1 << 10, 1 << 10, 4)])
def test_dev_commit(orc, shape):
    num_polys, poly_size, cw, depth = shape
    x = rand(5 + sum(shape), num_polys * poly_size)
    with W.Context() as ctx:
        dx = ctx.upload(x)
        cols = num_polys * depth
        dm = ctx.alloc(8 * cw * cols)
        dn = ctx.alloc(32 * W.merkle_num_nodes(cw))
        root = ctx.commit(dx, num_polys, poly_size, cw, depth, dm, dn)
        m = ctx.download(dm, cw * cols)
        nodes = ctx.download(dn, 32 * W.merkle_num_nodes(cw), np.uint8)
    want = orc.rs_encode(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(m, want)
    wn = orc.merkle_build(orc.commit_leaves_base(want, cw, cols), cw)
    assert np.array_equal(nodes, wn)
    assert root == wn[-32:].tobytes()


@pytest.mark.parametrize("poly_size,cw", [(1 << 12, 1 << 11), (1 << 15, 1 << 13), (1 << 16, 1 << 14),
                                          (1 << 16, 1 << 16)])
def test_dev_commit_ext3(orc, poly_size, cw):
    """Ext3 round commits: the 384-byte row leaves are hashed inside the
    encode (three 16-column sub-tiles per row) once the message length per
    column reaches 2048; every node of the tree against the oracle."""
    num_polys, depth = 1, 16
    x = rand(99 + poly_size + cw, 3 * poly_size)
    with W.Context() as ctx:
        dx = ctx.upload(x)
        dm = ctx.alloc(24 * cw * depth)
        dn = ctx.alloc(32 * W.merkle_num_nodes(cw))
        root = ctx.commit(dx, num_polys, poly_size, cw, depth, dm, dn, ext3=True)
        m = ctx.download(dm, 3 * cw * depth)
        nodes = ctx.download(dn, 32 * W.merkle_num_nodes(cw), np.uint8)
    want = orc.rs_encode_ext3(x, num_polys, poly_size, cw, depth)
    assert np.array_equal(m, want)
    wn = orc.merkle_build(orc.commit_leaves_ext3(want, cw, depth), cw)
    assert np.array_equal(nodes, wn)
    assert root == wn[-32:].tobytes()


# -------------------------------------------------------------- sumcheck

def lift(v):
    out = np.zeros(3 * v.size, dtype=np.uint64)
    out[0::3] = v
    return out


@pytest.mark.parametrize("n", [2, 4, 16, 1000, 1 << 14, 1 << 18])
def test_sumcheck_compute_fold(orc, n):
    a = rand(n, 3 * n)
    b = rand(n + 1, 3 * n)
    c0, c2 = W.sumcheck_compute(a, b)
    w0, w2 = orc.sumcheck_compute_ext3(a, b)
    assert np.array_equal(c0, w0) and np.array_equal(c2, w2)
    r = rand(n + 2, 3)
    assert np.array_equal(W.sumcheck_fold(a, r), orc.sumcheck_fold_ext3(a, r))
    f = rand(n + 3, n)
    c0, c2 = W.sumcheck_compute(f, b, a_is_base=True)
    w0, w2 = orc.sumcheck_compute_ext3(lift(f), b)
    assert np.array_equal(c0, w0) and np.array_equal(c2, w2)
    assert np.array_equal(W.sumcheck_fold(f, r, v_is_base=True), orc.sumcheck_fold_ext3(lift(f), r))


@pytest.mark.parametrize("n,a_base,inplace", [(2, True, False), (16, False, False), (1000, False, True),
                                              (1 << 16, True, False), (1 << 16, False, True), (1 << 20, True, False)])
def test_sumcheck_round_fused(orc, n, a_base, inplace):
    """compute + fold(a) + fold(b) in one sweep equals the three reference steps"""
    a = rand(3 * n + 1, n if a_base else 3 * n)
    b = rand(3 * n + 2, 3 * n)
    r = rand(3 * n + 3, 3)
    la = lift(a) if a_base else a
    w0, w2 = orc.sumcheck_compute_ext3(la, b)
    wa, wb = orc.sumcheck_fold_ext3(la, r), orc.sumcheck_fold_ext3(b, r)
    with W.Context() as ctx:
        da, db = ctx.upload(a), ctx.upload(b)
        if inplace:
            dfa, dfb = da, db
        else:
            dfa, dfb = ctx.alloc(24 * (n // 2)), ctx.alloc(24 * (n // 2))
        c0, c2 = ctx.sumcheck_round(da, a_base, db, n, r, dfa, dfb)
        fa, fb = ctx.download(dfa, 3 * (n // 2)), ctx.download(dfb, 3 * (n // 2))
    assert np.array_equal(c0, w0) and np.array_equal(c2, w2)
    assert np.array_equal(fa, wa) and np.array_equal(fb, wb)


def test_sumcheck_golden(fixtures):
    for c in fixtures["utils"]["sumcheck_compute"]:
        c0, c2 = W.sumcheck_compute(np.array(c["a"], dtype=np.uint64), lift(np.array(c["b"], dtype=np.uint64)),
                                    a_is_base=True)
        assert list(c0) == [c["acc0"], 0, 0] and list(c2) == [c["acc2"], 0, 0]
    f = fixtures["utils"]["fold"]
    out = W.sumcheck_fold(np.array(f["values"], dtype=np.uint64), [f["weight"], 0, 0], v_is_base=True)
    assert list(out[0::3]) == f["folded"]
    s = fixtures["sumcheck"]
    a, b = lift(np.array(s["a"], dtype=np.uint64)), lift(np.array(s["b"], dtype=np.uint64))
    for rd in s["rounds"]:
        c0, c2 = W.sumcheck_compute(a, b)
        assert c0[0] == rd["c0"] and c2[0] == rd["c2"]
        a = W.sumcheck_fold(a, [rd["r"], 0, 0])
        b = W.sumcheck_fold(b, [rd["r"], 0, 0])
    assert a[0] == s["final_a"] and b[0] == s["final_b"]


# ------------------------------------------------------------ OOD Horner

@pytest.mark.parametrize("n", [1, 2, 17, 256, 4097, 1 << 16, (1 << 20) + 3])
def test_ood_horner(orc, n):
    f = rand(n, n)
    pts = rand(n + 9, 6)  # two Ext3 points
    got = W.ood_evaluate_from_base(f, pts, 1, n)
    for p in range(2):
        assert np.array_equal(got[3 * p:3 * p + 3], orc.horner_base_ext3(f, pts[3 * p:3 * p + 3]))
    fe = rand(n + 5, 3 * n)
    got = W.ood_evaluate_from_ext3(fe, pts, 1, n)
    assert np.array_equal(got[:3], orc.horner_ext3_ext3(fe, pts[:3]))


def test_ood_horner_power_chunks(orc):
    """More coefficients than 256 power-table entries per lane (2^26 + 5 at
    2^18 lanes: 257 per lane): the sum runs over two chunks of the table, the
    second scaled by z^(256 T)."""
    n = (1 << 26) + 5
    f = rand(n, n)
    pts = rand(n + 9, 3)
    got = W.ood_evaluate_from_base(f, pts, 1, n)
    assert np.array_equal(got[:3], orc.horner_base_ext3(f, pts[:3]))


def test_ood_multi_vector_order(orc):
    nv, vs = 3, 1000
    f = rand(3, nv * vs)
    pts = rand(4, 6)
    got = W.ood_evaluate_from_base(f, pts, nv, vs)
    for p in range(2):
        for v in range(nv):
            want = orc.horner_base_ext3(f[v * vs:(v + 1) * vs], pts[3 * p:3 * p + 3])
            assert np.array_equal(got[3 * (p * nv + v):3 * (p * nv + v) + 3], want)


def test_ood_golden(fixtures):
    for c in fixtures["whir"]:
        vals = []
        for z in c["ood_points"]:
            for v in c["vectors"]:
                r = W.ood_evaluate_from_base(np.array(v, dtype=np.uint64), [z, 0, 0], 1, len(v))
                assert r[1] == 0 and r[2] == 0
                vals.append(int(r[0]))
        assert vals == c["ood_matrix"]
