"""Pin the CPU oracle against the reference's committed golden vectors.

Fixtures come from tests/golden/fixtures.json (made by make_fixtures.py from
the reference's cpp/dump_output goldens).
"""
import numpy as np
import pytest

P = 0xFFFFFFFF00000001


def test_base_field_golden(orc, fixtures):
    for c in fixtures["fields"]["base"]:
        a, b = orc.lib().or_gl_from_u64(c["a_raw"]), orc.lib().or_gl_from_u64(c["b_raw"])
        assert orc.gl_add(a, b) == c["add"]
        assert orc.gl_sub(a, b) == c["sub"]
        assert orc.lib().or_gl_neg(a) == c["neg"]
        assert orc.gl_mul(a, b) == c["mul"]
        assert orc.gl_mul(a, a) == c["sq"]
        assert orc.gl_inv(a) == c["inv"]
        assert orc.gl_pow(a, 10) == c["pow10"]


def test_ext3_field_golden(orc, fixtures):
    for c in fixtures["fields"]["ext3"]:
        a = [x % P for x in c["a_raw"]]
        b = [x % P for x in c["b_raw"]]
        assert list(orc.e3_add(a, b)) == c["add"]
        assert list(orc.e3_sub(a, b)) == c["sub"]
        assert list(orc.e3_mul(a, b)) == c["mul"]
        assert list(orc.e3_mul(a, a)) == c["sq"]
        assert list(orc.e3_inv(a)) == c["inv"]


def test_blake3_golden(orc, fixtures):
    n = 0
    for c in fixtures["hash"]:
        if c["engine"] != "blake3":
            continue
        data = bytes.fromhex(c["input"])
        for i in range(c["count"]):
            msg = data[i * c["size"]:(i + 1) * c["size"]]
            assert orc.blake3(msg).hex() == c["digests"][i]
            n += 1
    assert n == 9


def test_matrix_commit_golden(orc, fixtures):
    c0 = fixtures["matrix_commit"][0]
    leaves = orc.commit_leaves_base(np.array(c0["matrix"], dtype=np.uint64), c0["rows"], c0["cols"])
    got = [leaves[32 * i:32 * i + 32].tobytes().hex() for i in range(c0["rows"])]
    assert got == c0["leaves"]
    # Ext2 rows are c0,c1 LE per element: the same BLAKE3 over 128-byte rows
    c1 = fixtures["matrix_commit"][1]
    m = np.array(c1["matrix"], dtype=np.uint64)
    for r in range(c1["rows"]):
        row = m[r * 16:(r + 1) * 16].astype("<u8").tobytes()
        assert orc.blake3(row).hex() == c1["leaves"][r]


def test_merkle_golden(orc, fixtures):
    for c in fixtures["merkle_tree"]:
        if c["hash"] != "blake3":
            continue
        n = c["num_leaves"]
        leaves = np.frombuffer(bytes.fromhex(c["leaves"]), dtype=np.uint8)
        nodes = orc.merkle_build(leaves, n)
        assert nodes[-32:].tobytes().hex() == c["root"]
        hints = orc.merkle_open(nodes, n, c["indices"])
        assert [hints[32 * i:32 * i + 32].tobytes().hex() for i in range(len(hints) // 32)] == c["hints"]
        lh = np.concatenate([leaves[32 * i:32 * i + 32] for i in c["indices"]])
        assert orc.merkle_verify(bytes.fromhex(c["root"]), n, c["indices"], lh, hints)


def test_sha256_golden(orc, fixtures):
    """The SHA-256 engine's digests (dump_hash.cpp, engine "sha2")."""
    n = 0
    for c in fixtures["hash"]:
        if c["engine"] != "sha2":
            continue
        data = bytes.fromhex(c["input"])
        for i in range(c["count"]):
            assert orc.sha256(data[i * c["size"]:(i + 1) * c["size"]]).hex() == c["digests"][i]
            n += 1
    assert n > 0


def test_sha256_merkle_golden(orc, fixtures):
    """The SHA-256 tree (dump_merkle_tree.cpp "sha2-8-2"): root and open_path hints."""
    cases = [c for c in fixtures["merkle_tree"] if c["hash"] == "sha2"]
    assert cases
    for c in cases:
        n = c["num_leaves"]
        leaves = np.frombuffer(bytes.fromhex(c["leaves"]), dtype=np.uint8)
        nodes = orc.sha256_merkle_build(leaves, n)
        assert nodes[-32:].tobytes().hex() == c["root"]
        hints = orc.merkle_open(nodes, n, c["indices"])
        assert [hints[32 * i:32 * i + 32].tobytes().hex() for i in range(len(hints) // 32)] == c["hints"]


def test_sha256_commit_golden_whir(orc, fixtures):
    """golden_whir proofs (dump_whir.cpp) commit with SHA-256 trees: the
    narg string opens with the root of the witness matrix's row leaves."""
    for c in fixtures["whir"]:
        nv, vs = c["num_vectors"], c["vector_size"]
        m = np.array(c["witness_matrix"], dtype=np.uint64)
        nodes = orc.sha256_merkle_build(orc.sha256_leaves(m, 2 * vs, nv), 2 * vs)
        assert nodes[-32:].tobytes().hex() == c["proof_narg"][:64]


def test_sha256_oracle_vs_reference(orc):
    """Row leaves, trees (ragged sizes, zero padding) and hash_many against the
    reference's own Sha2 engine built from its sources."""
    R = orc.ref()
    if R is None:
        pytest.skip("reference oracle not built")
    rng = np.random.default_rng(7)
    for rows, cols in ((1, 1), (3, 7), (5, 8), (4, 9), (2, 16), (6, 15)):
        m = rng.integers(0, orc.P, size=rows * cols, dtype=np.uint64)
        ref = np.zeros(32 * rows, dtype=np.uint8)
        R.ref_commit_leaves_base_sha2(orc._ptr64(m), rows, cols, orc._ptr8(ref))
        assert (orc.sha256_leaves(m, rows, cols) == ref).all()
    for n in (1, 2, 3, 7, 17, 64):
        lv = rng.integers(0, 256, size=32 * n, dtype=np.uint8)
        ref = np.zeros(32 * orc.merkle_num_nodes(n), dtype=np.uint8)
        R.ref_merkle_build_sha2(orc._ptr8(lv), n, orc._ptr8(ref))
        assert (orc.sha256_merkle_build(lv, n) == ref).all()
    for size in (0, 1, 55, 56, 63, 64, 65, 119, 120, 200):
        data = rng.integers(0, 256, size=max(3 * size, 1), dtype=np.uint8)
        ref = np.zeros(96, dtype=np.uint8)
        R.ref_sha256_many(size, orc._ptr8(data), 3, orc._ptr8(ref))
        want = b"".join(orc.sha256(data[i * size:(i + 1) * size].tobytes()) for i in range(3))
        assert ref.tobytes() == want


def test_univariate_evaluate_golden(orc, fixtures):
    u = fixtures["utils"]["univariate_evaluate"]
    assert orc.horner_base(u["coeffs"], u["x"]) == u["result"]
    # the Ext3 Horner restricted to base-field points agrees
    r = orc.horner_base_ext3(u["coeffs"], [u["x"], 0, 0])
    assert list(r) == [u["result"], 0, 0]


def _lift(v):
    out = np.zeros(3 * len(v), dtype=np.uint64)
    out[0::3] = np.array(v, dtype=np.uint64)
    return out


def test_sumcheck_compute_golden(orc, fixtures):
    for c in fixtures["utils"]["sumcheck_compute"]:
        c0, c2 = orc.sumcheck_compute_ext3(_lift(c["a"]), _lift(c["b"]))
        assert list(c0) == [c["acc0"], 0, 0]
        assert list(c2) == [c["acc2"], 0, 0]
        c0b, c2b = orc.sumcheck_compute_base_ext3(c["a"], _lift(c["b"]))
        assert list(c0b) == list(c0) and list(c2b) == list(c2)


def test_fold_golden(orc, fixtures):
    f = fixtures["utils"]["fold"]
    out = orc.sumcheck_fold_ext3(_lift(f["values"]), [f["weight"], 0, 0])
    assert list(out[0::3]) == f["folded"]
    out2 = orc.sumcheck_fold_base_ext3(f["values"], [f["weight"], 0, 0])
    assert list(out2) == list(out)


def test_sumcheck_replay_golden(orc, fixtures):
    """Replay golden_sumcheck with its own r values: every c0, c2, sum and the
    final folded a, b (the claim algebra of sumcheck_protocol.hpp:108-140)."""
    s = fixtures["sumcheck"]
    a, b, total = _lift(s["a"]), _lift(s["b"]), [s["initial_sum"], 0, 0]
    for rd in s["rounds"]:
        c0, c2 = orc.sumcheck_compute_ext3(a, b)
        assert c0[0] == rd["c0"] and c2[0] == rd["c2"]
        c1 = orc.e3_sub(orc.e3_sub(total, orc.e3_add(c0, c0)), c2)
        r = [rd["r"], 0, 0]
        a = orc.sumcheck_fold_ext3(a, r)
        b = orc.sumcheck_fold_ext3(b, r)
        total = orc.e3_add(orc.e3_mul(orc.e3_add(orc.e3_mul(c2, r), c1), r), c0)
        assert total[0] == rd["sum"]
    assert a[0] == s["final_a"] and b[0] == s["final_b"]


def test_whir_witness_and_ood_golden(orc, fixtures):
    """golden_whir: RS encode (depth 1, codeword 2*size) and OOD Horner."""
    for c in fixtures["whir"]:
        nv, vs = c["num_vectors"], c["vector_size"]
        coeffs = np.array(sum(c["vectors"], []), dtype=np.uint64)
        m = orc.rs_encode(coeffs, nv, vs, 2 * vs, 1)
        assert list(m) == c["witness_matrix"]
        vals = [orc.horner_base(v, z) for z in c["ood_points"] for v in c["vectors"]]
        assert vals == c["ood_matrix"]
