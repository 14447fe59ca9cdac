"""GPU parity at the benchmark size: nv = 24 (2^24 coefficients), rate 1/4,
folding factor 4 -> 2^22 x 16 codeword, plus the first Ext3 sumcheck round.

Checks that do not need a full CPU recomputation:
  * sampled codeword rows against the closed form Y_b[i] = f_b(w_cw^i)
    (oracle Horner over each 2^20-coefficient block),
  * sampled leaves against oracle BLAKE3 of the downloaded rows, and their
    Merkle paths (GPU nodes) verified under the GPU root by the oracle,
  * linearity: LDE(f + g) = LDE(f) + LDE(g) on sampled rows,
  * the sumcheck round recomputed by the oracle in full (c0, c2, both folds),
and, when the reference build (oracle/_ref) is present, the complete
codeword and root against the reference's own CPU commit.
"""
import numpy as np
import pytest

import whirplonk_amd as W

pytestmark = pytest.mark.gpu

NV, DEPTH = 24, 16
POLY = 1 << NV
M = POLY // DEPTH
CW = 4 * M
P = 0xFFFFFFFF00000001


@pytest.fixture(scope="module")
def committed():
    assert W.device_available()
    coeffs, st = W.lcg_uniform(0x5748495200000002, POLY)
    ctx = W.Context(0)
    dx = ctx.upload(coeffs)
    dm = ctx.alloc(8 * CW * DEPTH)
    nn = W.merkle_num_nodes(CW)
    dn = ctx.alloc(32 * nn)
    root = ctx.commit(dx, 1, POLY, CW, DEPTH, dm, dn)
    matrix = ctx.download(dm, CW * DEPTH)
    nodes = ctx.download(dn, 32 * nn, np.uint8)
    yield {"ctx": ctx, "coeffs": coeffs, "state": st, "root": root, "matrix": matrix, "nodes": nodes, "dx": dx}
    ctx.close()


def test_sampled_rows_closed_form(orc, committed):
    coeffs, matrix = committed["coeffs"], committed["matrix"]
    w = orc.gl_root(CW)
    rng = np.random.default_rng(24)
    rows = [0, 1, CW - 1, CW // 2, CW // 4 + 3] + list(rng.integers(0, CW, 11))
    for i in rows:
        z = orc.gl_pow(w, int(i))
        for b in (0, 5, DEPTH - 1):
            want = orc.horner_base(coeffs[b * M:(b + 1) * M], z)
            assert int(matrix[i * DEPTH + b]) == want, (i, b)


def test_leaves_and_paths(orc, committed):
    matrix, nodes, root = committed["matrix"], committed["nodes"], committed["root"]
    assert nodes[-32:].tobytes() == root
    rng = np.random.default_rng(5)
    idx = [0, CW - 1] + list(rng.integers(0, CW, 62)) + [7, 7]
    leaves = orc.commit_leaves_base(np.concatenate([matrix[i * DEPTH:(i + 1) * DEPTH] for i in idx]), len(idx), DEPTH)
    for k, i in enumerate(idx):
        assert leaves[32 * k:32 * k + 32].tobytes() == nodes[32 * i:32 * i + 32].tobytes()
    ni = W.merkle_hint_node_indices(CW, idx)
    hints = np.concatenate([nodes[32 * int(j):32 * int(j) + 32] for j in ni])
    assert orc.merkle_verify(root, CW, idx, leaves, hints)
    bad = leaves.copy()
    bad[0] ^= 1
    assert not orc.merkle_verify(root, CW, idx, bad, hints)


def test_linearity(committed):
    ctx, coeffs, matrix = committed["ctx"], committed["coeffs"], committed["matrix"]
    g, _ = W.lcg_uniform(99, POLY)
    s = coeffs.astype(object) + g.astype(object)
    fg = np.array([int(v) % P for v in s[: 1 << 16]] + [0], dtype=np.uint64)  # only used below on sampled rows
    # full f+g on the host with numpy (mod p without overflow: both < p < 2^64)
    with np.errstate(over="ignore"):
        t = coeffs + g
        wrapped = t < coeffs
        t = np.where(wrapped | (t >= np.uint64(P)), t - np.uint64(P), t)
    assert np.array_equal(t[: 1 << 16], fg[: 1 << 16])
    dg = ctx.upload(g)
    dfg = ctx.upload(t)
    dmg = ctx.alloc(8 * CW * DEPTH)
    dmfg = ctx.alloc(8 * CW * DEPTH)
    ctx.rs_encode(dg, 1, POLY, CW, DEPTH, dmg)
    ctx.rs_encode(dfg, 1, POLY, CW, DEPTH, dmfg)
    mg = ctx.download(dmg, CW * DEPTH)
    mfg = ctx.download(dmfg, CW * DEPTH)
    with np.errstate(over="ignore"):
        u = matrix + mg
        u = np.where((u < matrix) | (u >= np.uint64(P)), u - np.uint64(P), u)
    assert np.array_equal(u, mfg)
    for p_ in (dg, dfg, dmg, dmfg):
        ctx.free(p_)


def test_sumcheck_round_full(orc, committed):
    ctx, coeffs, st = committed["ctx"], committed["coeffs"], committed["state"]
    b, st = W.lcg_uniform(st, 3 * POLY)
    r, _ = W.lcg_uniform(st, 3)
    db = ctx.upload(b)
    dfa, dfb = ctx.alloc(24 * (POLY // 2)), ctx.alloc(24 * (POLY // 2))
    c0, c2 = ctx.sumcheck_round(committed["dx"], True, db, POLY, r, dfa, dfb)
    fa, fb = ctx.download(dfa, 3 * (POLY // 2)), ctx.download(dfb, 3 * (POLY // 2))
    w0, w2 = orc.sumcheck_compute_base_ext3(coeffs, b)
    assert np.array_equal(c0, w0) and np.array_equal(c2, w2)
    assert np.array_equal(fa, orc.sumcheck_fold_base_ext3(coeffs, r))
    assert np.array_equal(fb, orc.sumcheck_fold_ext3(b, r))
    for p_ in (db, dfa, dfb):
        ctx.free(p_)


def test_against_reference_build(orc, committed):
    R = orc.ref()
    if R is None:
        pytest.skip("oracle/_ref (the reference's own CPU build) is not present on this machine")
    coeffs = committed["coeffs"]
    out = np.empty(CW * DEPTH, dtype=np.uint64)
    root = np.zeros(32, dtype=np.uint8)
    R.ref_commit_base(orc._ptr64(coeffs), POLY, CW, DEPTH, orc._ptr64(out), orc._ptr8(root))
    assert np.array_equal(out, committed["matrix"])
    assert root.tobytes() == committed["root"]
