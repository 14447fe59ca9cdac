"""The host-pointer drop-ins that the reference's NTT engine calls, at its
own call sites: gpu_ntt_batch_transpose (cuda_integration.hpp:388-430), which
ark_ntt offers the zero-padded interleaved matrix (mod_ntt.hpp:77-98), against
the reference build's interleaved_rs_encode / ntt_batch and the oracle."""
import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu


def _padded(coeffs, num_polys, poly_size, cw, depth):
    """ark_ntt's step 1 (mod_ntt.hpp:77-93): block b of poly k at
    k*cw*depth + b*cw, its m coefficients first, zeros after."""
    m = poly_size // depth
    out = np.zeros(num_polys * depth * cw, dtype=np.uint64)
    for k in range(num_polys):
        for b in range(depth):
            dst = (k * depth + b) * cw
            out[dst:dst + m] = coeffs[k * poly_size + b * m:k * poly_size + (b + 1) * m]
    return out


def _rs_ref(coeffs, num_polys, poly_size, cw, depth):
    R = oracle.ref()
    if R is None:
        return oracle.rs_encode(coeffs, num_polys, poly_size, cw, depth)
    out = np.zeros(cw * num_polys * depth, dtype=np.uint64)
    R.ref_rs_encode(oracle._ptr64(coeffs), num_polys, poly_size, cw, depth, oracle._ptr64(out))
    return out


@pytest.mark.parametrize("num_polys,poly_size,cw,depth", [
    (1, 1 << 16, 1 << 14, 16), (3, 1 << 14, 1 << 12, 16), (2, 1 << 12, 1 << 13, 4), (1, 1 << 21, 1 << 21, 16)])
def test_ntt_batch_transpose_ark_ntt_call_site(num_polys, poly_size, cw, depth):
    """values = ark_ntt's padded matrix, ntt_size = cols = codeword_length,
    rows = num_polys * depth: the result is the interleaved RS codeword."""
    coeffs, _ = W.lcg_uniform(poly_size + cw + num_polys, num_polys * poly_size)
    vals = _padded(coeffs, num_polys, poly_size, cw, depth)
    print("ntt transpose call site", num_polys, poly_size, cw, depth, flush=True)
    got = W.ntt_batch_transpose(vals, cw, num_polys * depth, cw)
    assert np.array_equal(got, _rs_ref(coeffs, num_polys, poly_size, cw, depth))


@pytest.mark.parametrize("total,ntt_size,rows,cols", [
    (1 << 16, 1 << 10, 1 << 5, 1 << 11), (3 << 15, 1 << 10, 3, 1 << 15), (48 << 12, 1 << 12, 48, 1 << 12),
    (1 << 22, 1 << 21, 2, 1 << 21), (1 << 17, 1 << 17, 1 << 7, 1 << 10)])
def test_ntt_batch_transpose_shapes(total, ntt_size, rows, cols):
    """Any rows x cols split of the NTT'd blocks, as the reference's
    transpose_kernel (cuda_ntt.cu:149-162) allows: the blocks' forward NTT
    (the reference build's ntt_batch, or the oracle), then the transpose."""
    vals, _ = W.lcg_uniform(total + rows, total)
    R = oracle.ref()
    if R is not None:
        want = vals.copy()
        R.ref_ntt_batch(oracle._ptr64(want), total, ntt_size)
    else:
        want = oracle.ntt_batch(vals, ntt_size)
    want = want.reshape(rows, cols).T.reshape(-1)
    print("ntt transpose", total, ntt_size, rows, cols, flush=True)
    assert np.array_equal(W.ntt_batch_transpose(vals, ntt_size, rows, cols), want)
    # device form: out of place, input untouched
    with W.Context() as ctx:
        d_in = ctx.upload(vals)
        d_out = ctx.alloc(8 * total)
        ctx.ntt_batch_transpose(d_in, d_out, total, ntt_size, rows, cols)
        assert np.array_equal(ctx.download(d_out, total), want)
        assert np.array_equal(ctx.download(d_in, total), vals)


def test_ntt_batch_transpose_gates():
    lib = W.load()
    vals, _ = W.lcg_uniform(9, 1 << 12)
    v = vals.copy()
    # below the threshold (65536 elements): not handled, values untouched
    assert lib.whir_gpu_ntt_batch_transpose(W._p64(v), v.size, 1 << 10, 4, 1 << 10) == 1
    assert np.array_equal(v, vals)
    t = lib.whir_gpu_ntt_threshold()
    lib.whir_set_gpu_ntt_threshold(1 << 12)
    try:
        want = oracle.ntt_batch(vals, 1 << 10).reshape(4, 1 << 10).T.reshape(-1)
        assert np.array_equal(W.ntt_batch_transpose(vals, 1 << 10, 4, 1 << 10), want)
        # inconsistent shapes are refused
        assert lib.whir_gpu_ntt_batch_transpose(W._p64(v), v.size, 1 << 10, 3, 1 << 10) > 0
        assert lib.whir_gpu_ntt_batch_transpose(W._p64(v), v.size, 3 << 8, 4, 1 << 10) > 0
    finally:
        lib.whir_set_gpu_ntt_threshold(t)


def test_integration_call_sites_compiled_against_reference():
    """oracle/_ref/integration_check: every call-site edit of the integration
    notes, compiled against the reference's own headers and types (std::span,
    std::vector<Goldilocks>, Hash) and linked to libwhir_gpu.so, gives the
    reference CPU function's result (ntt_batch, intt_batch, ark_ntt's
    transpose, interleaved_rs_encode, the encode-and-leaves step in base and
    Ext3 form, BLAKE3 and SHA-256 trees and open paths, OOD evaluation,
    encode_into, hash_many, compute_sumcheck_polynomial and fold)."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(oracle.__file__), "_ref", "integration_check")
    if not os.path.exists(exe):
        pytest.skip("integration driver not built (needs the reference headers)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok  ") >= 18 and "FAIL" not in r.stdout


def test_whir_profile_csv_rows():
    """WHIR_PROFILE=1 makes the drop-ins and the prover print the reference's
    CSV rows (profiling.hpp:12-142): one `mode,phase,size,stage,time_ms`
    header, then rows whose stages are the reference's kept vocabulary, with
    the commit/prove phases of its PhaseGuard."""
    import os
    import subprocess
    import sys

    code = (
        "import numpy as np, whirplonk_amd as W\n"
        "c, st = W.lcg_uniform(3, 1 << 14)\n"
        "W.rs_encode_leaves(c, 1, 1 << 14, 1 << 12, 16)\n"
        "pts, _ = W.lcg_uniform(st, 3 * 14)\n"
        "pts = pts.reshape(1, 14, 3)\n"
        "p = W.Params.default(starting_log_inv_rate=2)\n"
        "pid, sid = W.domain_separator_cli(p)\n"
        "with W.Context() as ctx:\n"
        "    d = ctx.upload(c)\n"
        "    ev = np.array([ctx.mle_evaluate(d, True, pts[0])])\n"
        "    with W.Prover(ctx, p, 1 << 14, pid, sid) as pr:\n"
        "        pr.commit(d, 1)\n"
        "        pr.prove(pts, ev)\n"
    )
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, WHIR_PROFILE="1", PYTHONPATH=root)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env, cwd=root)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stderr.splitlines() if l.count(",") == 4]
    assert lines[0] == "mode,phase,size,stage,time_ms" and lines.count(lines[0]) == 1
    rows = [l.split(",") for l in lines[1:]]
    kept = {"witness_encoding", "witness_gpu_alloc", "witness_h2d", "witness_rs_encode", "witness_leaf_hash",
            "witness_d2h", "merkle_build_total", "ood_evaluation", "commit_total", "prove_total", "total_prover"}
    assert rows and all(st in kept and float(ms) >= 0 and int(sz) > 0 for _, _, sz, st, ms in rows)
    got = {(m, ph, st) for m, ph, _, st, _ in rows}
    for want in [("gpu", "none", "witness_h2d"), ("gpu", "none", "witness_rs_encode"), ("gpu", "none", "witness_d2h"),
                 ("prover", "commit", "witness_encoding"), ("prover", "commit", "merkle_build_total"),
                 ("prover", "commit", "ood_evaluation"), ("prover", "commit", "commit_total"),
                 ("prover", "prove", "prove_total"), ("prover", "none", "total_prover")]:
        assert want in got, want
    # without the variable nothing is printed
    env.pop("WHIR_PROFILE")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env, cwd=root)
    assert r.returncode == 0 and "mode,phase,size,stage,time_ms" not in r.stderr


@pytest.mark.parametrize("nbytes", [(4 << 20) - 8, 4 << 20, (16 << 20) + 8, (37 << 20) + 24, 96 << 20])
def test_staged_host_copies(nbytes):
    """Copies between pageable host memory and the device go through two
    pinned staging chunks (16 MiB) above 4 MiB: sizes below, at and between
    the chunk boundaries come back unchanged, from pageable and from pinned
    host memory, and a device-to-device check sees the uploaded bytes."""
    n = nbytes // 8
    src, _ = W.lcg_uniform(nbytes, n)
    print("staged copy", nbytes, flush=True)
    with W.Context() as ctx:
        d = ctx.upload(src)
        assert np.array_equal(ctx.download(d, n), src)
        # the encode-to-bytes drop-in: host -> device -> host through the staged paths
        out = np.empty(nbytes, dtype=np.uint8)
        W._check(W.load().whir_gpu_encode_to_bytes(W._p64(src), W._p8(out), n))
        assert np.array_equal(out.view(np.uint64), src)
        pin = W.Pinned(n)
        pin.array[:] = src[::-1]
        W._check(W.load().whir_memcpy_h2d(ctx.h, d, pin.ptr, nbytes))
        pin.array[:] = 0
        W._check(W.load().whir_memcpy_d2h(ctx.h, pin.ptr, d, nbytes))
        assert np.array_equal(pin.array, src[::-1])
        assert np.array_equal(ctx.download(d, n), src[::-1])
        ctx.free(d)


def test_staged_copies_from_threads():
    """Several host threads copying through one context at once: the copies
    share the context's staging chunks, so each takes the context's lock and
    every buffer still arrives and returns intact."""
    import threading

    nthreads, n = 4, (40 << 20) // 8
    srcs = [W.lcg_uniform(1234 + t, n)[0] for t in range(nthreads)]
    errors = []
    print("staged copies from", nthreads, "threads", flush=True)
    with W.Context() as ctx:
        bufs = [ctx.alloc(8 * n) for _ in range(nthreads)]

        def work(t):
            try:
                for _ in range(3):
                    W._check(W.load().whir_memcpy_h2d(ctx.h, bufs[t], W._p64(srcs[t]), 8 * n))
                    back = np.empty(n, dtype=np.uint64)
                    W._check(W.load().whir_memcpy_d2h(ctx.h, W._p64(back), bufs[t], 8 * n))
                    if not np.array_equal(back, srcs[t]):
                        errors.append(f"thread {t}: data changed")
            except Exception as e:  # reported below, on the test's thread
                errors.append(f"thread {t}: {e}")

        threads = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        for b in bufs:
            ctx.free(b)
    assert not errors, errors
