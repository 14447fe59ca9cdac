"""Pin the oracle where the goldens leave gaps: closed forms, brute force,
invariants, and the reference's own C++ build (oracle/_ref) when present.

Each check is chosen so that a plausible oracle mistake fails it: a wrong
root (closed-form DFT), a missing bit reversal (DFT), 1/N scaling in the
inverse (NTT(INTT(x)) = N x), a transposed RS layout (closed form per
row/column), a wrong zero-hash padding (brute-force tree), a wrong Ext3
non-residue (polynomial product mod X^3 - 2 in big ints).
"""
import random

import numpy as np
import pytest

P = 0xFFFFFFFF00000001
SPECIAL = [0, 1, P - 1, P - 2, 2**32 - 1, 2**32, 7, 2**63, P - 2**32]


def test_field_bigint(orc):
    rnd = random.Random(1)
    vals = SPECIAL + [rnd.randrange(P) for _ in range(200)]
    for a in vals:
        for b in vals[:20]:
            assert orc.gl_add(a, b) == (a + b) % P
            assert orc.gl_sub(a, b) == (a - b) % P
            assert orc.gl_mul(a, b) == (a * b) % P
    assert orc.lib().or_gl_from_u64(2**64 - 1) == (2**64 - 1) - P


def _e3mul_big(a, b):
    # polynomial product, then X^3 = 2, X^4 = 2X
    c = [0] * 5
    for i in range(3):
        for j in range(3):
            c[i + j] += a[i] * b[j]
    return [(c[0] + 2 * c[3]) % P, (c[1] + 2 * c[4]) % P, c[2] % P]


def test_ext3_bigint(orc):
    rnd = random.Random(2)
    for _ in range(200):
        a = [rnd.choice(SPECIAL + [rnd.randrange(P)]) for _ in range(3)]
        b = [rnd.randrange(P) for _ in range(3)]
        assert list(orc.e3_mul(a, b)) == _e3mul_big(a, b)
        if any(a):
            assert list(orc.e3_mul(a, orc.e3_inv(a))) == [1, 0, 0]


def test_roots(orc):
    w32 = pow(7, 2**32 - 1, P)
    for k in range(0, 33):
        n = 1 << k
        w = orc.gl_root(n)
        assert w == pow(w32, 2**32 // n, P)
        assert pow(w, n, P) == 1
        if n > 1:
            assert pow(w, n // 2, P) == P - 1
    # the shift-friendly small roots used by the GPU stages
    assert orc.gl_root(64) == pow(2, 39, P)
    assert orc.gl_root(8) == pow(2, 120, P)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 256, 1024])
def test_ntt_vs_bruteforce(orc, n):
    rnd = random.Random(n)
    x = np.array([rnd.choice(SPECIAL) if rnd.random() < 0.2 else rnd.randrange(P) for _ in range(n)],
                 dtype=np.uint64)
    got = orc.ntt_batch(x, n)
    assert np.array_equal(got, orc.dft_naive(x))
    # closed form for a few outputs in Python big ints
    w = orc.gl_root(n)
    for k in sorted({0, min(1, n - 1), n // 2, n - 1}):
        assert int(got[k]) == sum(int(x[j]) * pow(w, j * k, P) for j in range(n)) % P


def test_intt_unscaled(orc):
    rnd = random.Random(7)
    for n in (2, 8, 4096):
        x = np.array([rnd.randrange(P) for _ in range(3 * n)], dtype=np.uint64)
        y = orc.intt_batch(orc.ntt_batch(x, n), n)
        assert np.array_equal(y, np.array([(int(v) * n) % P for v in x], dtype=np.uint64))


def test_rs_closed_form(orc):
    rnd = random.Random(3)
    num_polys, poly_size, depth, cw = 2, 32, 4, 32
    m = poly_size // depth
    coeffs = np.array([rnd.randrange(P) for _ in range(num_polys * poly_size)], dtype=np.uint64)
    out = orc.rs_encode(coeffs, num_polys, poly_size, cw, depth)
    cols = num_polys * depth
    w = orc.gl_root(cw)
    for i in (0, 1, 5, cw - 1):
        for k in range(num_polys):
            for b in range(depth):
                f = [int(coeffs[k * poly_size + b * m + j]) for j in range(m)]
                want = sum(f[j] * pow(w, i * j, P) for j in range(m)) % P
                assert int(out[i * cols + k * depth + b]) == want


def test_rs_ext3_componentwise(orc):
    rnd = random.Random(4)
    n = 64
    c = np.array([rnd.randrange(P) for _ in range(3 * n)], dtype=np.uint64)
    out = orc.rs_encode_ext3(c, 1, n, 4 * n // 4, 4)
    for comp in range(3):
        assert np.array_equal(out[comp::3], orc.rs_encode(c[comp::3], 1, n, n, 4))


def test_merkle_bruteforce(orc):
    rnd = random.Random(5)
    for n in (0, 1, 2, 3, 17, 64):
        leaves = np.array([rnd.randrange(256) for _ in range(32 * n)], dtype=np.uint8)
        nodes = orc.merkle_build(leaves, n)
        # recursive root with zero padding to the next power of two
        width = 1
        while width < n:
            width *= 2
        layer = [leaves[32 * i:32 * i + 32].tobytes() for i in range(n)] + [bytes(32)] * (width - n)
        while len(layer) > 1:
            layer = [orc.blake3(layer[2 * i] + layer[2 * i + 1]) for i in range(len(layer) // 2)]
        assert nodes[-32:].tobytes() == layer[0]
        if n:
            idx = [rnd.randrange(n) for _ in range(5)] + [0]
            hints = orc.merkle_open(nodes, n, idx)
            lh = np.concatenate([leaves[32 * i:32 * i + 32] for i in idx])
            assert orc.merkle_verify(layer[0], n, idx, lh, hints)
            bad = hints.copy()
            if bad.size:
                bad[0] ^= 1
                assert not orc.merkle_verify(layer[0], n, idx, lh, bad)


def test_blake3_multichunk_structure(orc):
    """>1024-byte messages: chunk CVs merged by parent nodes.  Values are pinned
    against the reference build (TestAgainstReferenceBuild.test_blake3_all_lengths);
    this CPU-only check keeps length sensitivity across the chunk boundaries."""
    d = bytes(range(256)) * 9
    h = {n: orc.blake3(d[:n]) for n in (1023, 1024, 1025, 2048, 2049, 2304)}
    assert len(set(h.values())) == len(h)


@pytest.mark.skipif("not __import__('oracle').ref()")
class TestAgainstReferenceBuild:
    """The reference's own C++ (compiled by oracle/Makefile) on the same inputs."""

    def test_ntt(self, orc):
        R = orc.ref()
        for n in (2, 64, 4096):
            x, _ = orc.lcg(11 + n, 3 * n)
            y = x.copy()
            R.ref_ntt_batch(orc._ptr64(y), y.size, n)
            assert np.array_equal(y, orc.ntt_batch(x, n))

    @pytest.mark.parametrize("shape", [(1, 64, 256, 16), (2, 256, 1024, 4), (1, 4096, 16384, 16), (3, 1024, 512, 4)])
    def test_rs_encode(self, orc, shape):
        R = orc.ref()
        num_polys, poly_size, cw, depth = shape
        x, _ = orc.lcg(99, num_polys * poly_size)
        out = np.zeros(cw * num_polys * depth, dtype=np.uint64)
        R.ref_rs_encode(orc._ptr64(x), num_polys, poly_size, cw, depth, orc._ptr64(out))
        assert np.array_equal(out, orc.rs_encode(x, num_polys, poly_size, cw, depth))

    def test_rs_encode_ext3(self, orc):
        R = orc.ref()
        x, _ = orc.lcg(5, 3 * 256)
        out = np.zeros(3 * 256 * 4 // 4 * 4, dtype=np.uint64)
        R.ref_rs_encode_ext3(orc._ptr64(x), 1, 256, 256, 4, orc._ptr64(out))
        assert np.array_equal(out, orc.rs_encode_ext3(x, 1, 256, 256, 4))

    def test_leaves_and_tree(self, orc):
        R = orc.ref()
        for rows, cols in ((1000, 16), (64, 5), (33, 48)):
            m, _ = orc.lcg(rows * cols, rows * cols)
            a = np.zeros(32 * rows, dtype=np.uint8)
            R.ref_commit_leaves_base(orc._ptr64(m), rows, cols, orc._ptr8(a))
            assert np.array_equal(a, orc.commit_leaves_base(m, rows, cols))
            nodes = np.zeros(32 * orc.merkle_num_nodes(rows), dtype=np.uint8)
            R.ref_merkle_build(orc._ptr8(a), rows, orc._ptr8(nodes))
            assert np.array_equal(nodes, orc.merkle_build(a, rows))
            idx = np.array([3, 3, 0, rows - 1, 17], dtype=np.uint64)
            nh = R.ref_merkle_open(orc._ptr8(nodes), rows, orc._ptr64(idx), idx.size, None)
            h = np.zeros(32 * nh, dtype=np.uint8)
            R.ref_merkle_open(orc._ptr8(nodes), rows, orc._ptr64(idx), idx.size, orc._ptr8(h))
            assert np.array_equal(h, orc.merkle_open(nodes, rows, list(idx)))
        m, _ = orc.lcg(1, 3 * 16 * 40)
        a = np.zeros(32 * 40, dtype=np.uint8)
        R.ref_commit_leaves_ext3(orc._ptr64(m), 40, 16, orc._ptr8(a))
        assert np.array_equal(a, orc.commit_leaves_ext3(m, 40, 16))

    @pytest.mark.parametrize("size", [0, 1, 63, 64, 100, 1024, 1025, 2048, 3072, 4097, 9000])
    def test_blake3_all_lengths(self, orc, size):
        """hash_many's SIMD branch (<= 1024, multiple of 64) and its incremental
        hasher branch (any other size, multi-chunk trees) against the C restatement"""
        R = orc.ref()
        count = 3
        data = np.random.default_rng(size).integers(0, 256, size * count + 1, dtype=np.uint8)
        out = np.zeros(32 * count, dtype=np.uint8)
        R.ref_blake3_many(size, orc._ptr8(data), count, orc._ptr8(out))
        for i in range(count):
            assert out[32 * i:32 * i + 32].tobytes() == orc.blake3(data[i * size:(i + 1) * size].tobytes())

    def test_sumcheck_and_horner(self, orc):
        R = orc.ref()
        n = 1 << 13
        a, st = orc.lcg(21, 3 * n)
        b, st = orc.lcg(st, 3 * n)
        c0, c2 = np.zeros(3, dtype=np.uint64), np.zeros(3, dtype=np.uint64)
        R.ref_sumcheck_compute_ext3(orc._ptr64(a), orc._ptr64(b), n, orc._ptr64(c0), orc._ptr64(c2))
        o0, o2 = orc.sumcheck_compute_ext3(a, b)
        assert np.array_equal(c0, o0) and np.array_equal(c2, o2)
        r = np.array([5, 6, 7], dtype=np.uint64)
        v = a.copy()
        R.ref_sumcheck_fold_ext3(orc._ptr64(v), n, orc._ptr64(r))
        assert np.array_equal(v[: 3 * n // 2], orc.sumcheck_fold_ext3(a, r))
        z = np.array([123456789, 987654321, 555], dtype=np.uint64)
        out = np.zeros(3, dtype=np.uint64)
        R.ref_horner_base_ext3(orc._ptr64(b[:n]), n, orc._ptr64(z), orc._ptr64(out))
        assert np.array_equal(out, orc.horner_base_ext3(b[:n], z))


def _e3pow_big(a, e):
    # square-and-multiply over big-int polynomial products mod X^3 - 2
    r, x = [1, 0, 0], list(a)
    while e:
        if e & 1:
            r = _e3mul_big(r, x)
        x = _e3mul_big(x, x)
        e >>= 1
    return r


def test_ext3_pow_bigint(orc):
    """or_e3_pow against big-int square-and-multiply (exponents 0, 1, 2, the
    multiplicative group order minus one, so a^-1, and random 64-bit ones),
    and against GoldilocksExt3::pow of the reference build when present."""
    rnd = random.Random(4)
    order = P**3 - 1
    R = orc.ref()
    for _ in range(60):
        a = [rnd.choice(SPECIAL + [rnd.randrange(P)]) for _ in range(3)]
        for e in (0, 1, 2, 3, P - 1, P, 2**64 - 1, rnd.randrange(2**64)):
            assert list(orc.e3_pow(a, e)) == _e3pow_big(a, e), (a, e)
            if R is not None:
                got = np.zeros(3, dtype=np.uint64)
                R.ref_e3_pow(orc._ptr64(orc.u64(a)), e, orc._ptr64(got))
                assert list(orc.e3_pow(a, e)) == [int(x) for x in got], (a, e)
        if any(a):
            # a^(p^3 - 2) = a^-1 (the exponent exceeds 64 bits: two steps)
            inv = _e3pow_big(a, order - 1)
            assert _e3mul_big(a, inv) == [1, 0, 0]
            assert list(orc.e3_inv(a)) == inv


def test_horner_ext3_vs_reference(orc):
    """or_horner_ext3_ext3 against the reference's mixed_univariate_evaluate
    over Ext3 coefficients (Identity embedding, utilities.hpp:195-246), both
    its serial and its chunked OpenMP branch (>= 4096 coefficients), and
    against a big-int sum of c_i z^i on small inputs."""
    R = orc.ref()
    rnd = random.Random(5)
    for n in (1, 2, 3, 7, 64, 1000):
        c = [rnd.randrange(P) for _ in range(3 * n)]
        z = [rnd.randrange(P) for _ in range(3)]
        want, zi = [0, 0, 0], [1, 0, 0]
        for i in range(n):
            t = _e3mul_big(c[3 * i:3 * i + 3], zi)
            want = [(want[k] + t[k]) % P for k in range(3)]
            zi = _e3mul_big(zi, z)
        assert list(orc.horner_ext3_ext3(c, z)) == want, n
    if R is None:
        pytest.skip("reference build not available")
    for n in (1, 5, 4095, 4096, 4097, 20000):
        c = orc.u64([rnd.randrange(P) for _ in range(3 * n)])
        z = orc.u64([rnd.randrange(P) for _ in range(3)])
        got = np.zeros(3, dtype=np.uint64)
        R.ref_horner_ext3_ext3(orc._ptr64(c), n, orc._ptr64(z), orc._ptr64(got))
        assert np.array_equal(orc.horner_ext3_ext3(c, z), got), n


def test_pow_find_vs_reference(orc):
    """The serial PoW scans (BLAKE3 in C, SHA-256 with hashlib) against the
    reference's find_nonce (proof_of_work.hpp:95-193) with either engine, at
    thresholds from zero difficulty to about 12 bits."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference build not available")
    rnd = random.Random(6)
    for trial in range(6):
        ch = bytes(rnd.randrange(256) for _ in range(32))
        for bits in (0, 3, 8, 12):
            thr = 2**64 - 1 if bits == 0 else int(2 ** (64 - bits))
            chb = np.frombuffer(ch, dtype=np.uint8).copy()
            want_b3 = int(R.ref_pow_find(1, orc._ptr8(chb), thr))
            want_sha = int(R.ref_pow_find(2, orc._ptr8(chb), thr))
            assert orc.pow_find(ch, thr) == want_b3, (trial, bits)
            assert orc.sha256_pow_find(ch, thr) == want_sha, (trial, bits)
