"""whirplonk_amd — MI355X (gfx950) WHIR commit-and-fold path.

Python view of the C ABI in include/whir_gpu.h (libwhir_gpu.so, built
in-tree by ``whirplonk_amd.build``).  The library is the product; this module
only loads it with ctypes and converts numpy arrays.  Importing never falls
back to anything: if the shared library is missing or fails to load, every
call raises.

Mirrors the reference's GPU entry points (cpp/cuda/cuda_integration.hpp) and
its CPU algebra calls (ntt_batch, interleaved_rs_encode, commit_leaves,
merkle build/open, compute_sumcheck_polynomial, fold,
mixed_univariate_evaluate) with the same argument meaning.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libwhir_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "whir_gpu.h")

P = 0xFFFFFFFF00000001

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_int = ctypes.c_int


class WhirGpuError(RuntimeError):
    """A device error (status < 0) or an unsupported request (status > 0)."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


# (name, restype, argtypes) for every exported symbol of include/whir_gpu.h
_SIGS = [
    ("whir_last_error", ctypes.c_char_p, []),
    ("whir_device_available", _int, []),
    ("whir_gpu_dispatch_enabled", _int, []),
    ("whir_set_gpu_dispatch_enabled", None, [_int]),
    ("whir_gpu_ntt_threshold", _sz, []),
    ("whir_set_gpu_ntt_threshold", None, [_sz]),
    ("whir_gpu_warmup", _int, []),
    ("whir_ctx_create", _int, [_int, ctypes.POINTER(_vp)]),
    ("whir_ctx_destroy", _int, [_vp]),
    ("whir_ctx_set_stream", _int, [_vp, _vp]),
    ("whir_ctx_set_hash", _int, [_vp, ctypes.c_uint32]),
    ("whir_ctx_stream", _vp, [_vp]),
    ("whir_ctx_synchronize", _int, [_vp]),
    ("whir_dev_malloc", _int, [_vp, _sz, ctypes.POINTER(_vp)]),
    ("whir_dev_free", _int, [_vp, _vp]),
    ("whir_memcpy_h2d", _int, [_vp, _vp, _vp, _sz]),
    ("whir_memcpy_d2h", _int, [_vp, _vp, _vp, _sz]),
    ("whir_dev_copy", _int, [_vp, _vp, _vp, _sz]),
    ("whir_gpu_ntt_batch", _int, [_u64p, _sz, _sz]),
    ("whir_gpu_intt_batch", _int, [_u64p, _sz, _sz]),
    ("whir_gpu_ntt_batch_transpose", _int, [_u64p, _sz, _sz, _sz, _sz]),
    ("whir_gpu_interleaved_rs_encode", _int, [_u64p, _u64p, _sz, _sz, _sz, _sz]),
    ("whir_gpu_interleaved_rs_encode_blake3_matrix_leaves", _int, [_u64p, _u64p, _u8p, _sz, _sz, _sz, _sz]),
    ("whir_gpu_interleaved_rs_encode_blake3_ext3_matrix_leaves", _int,
     [_u64p, _u64p, _u64p, _u64p, _u64p, _u64p, _u8p, _sz, _sz, _sz, _sz]),
    ("whir_gpu_blake3_merkle_tree", _int, [_u8p, _sz, _u8p, _u8p]),
    ("whir_gpu_sha256_merkle_tree", _int, [_u8p, _sz, _u8p, _u8p]),
    ("whir_merkle_hint_node_indices", _sz, [_sz, _u64p, _sz, _u64p]),
    ("whir_gpu_blake3_merkle_open_path", _int, [_u8p, _sz, _u64p, _sz, _u8p, _u8p]),
    ("whir_gpu_sha256_merkle_open_path", _int, [_u8p, _sz, _u64p, _sz, _u8p, _u8p]),
    ("whir_gpu_ood_evaluate_ext3_from_base", _int, [_u64p, _u64p, _u64p, _sz, _sz, _sz]),
    ("whir_gpu_ood_evaluate_ext3_from_ext3", _int, [_u64p, _u64p, _u64p, _u64p, _u64p, _sz, _sz, _sz]),
    ("whir_gpu_encode_to_bytes", _int, [_u64p, _u8p, _sz]),
    ("whir_gpu_blake3_hash_many", _int, [_sz, _u8p, _sz, _u8p]),
    ("whir_gpu_sha256_hash_many", _int, [_sz, _u8p, _sz, _u8p]),
    ("whir_gpu_sumcheck_compute", _int, [_u64p, _int, _u64p, _sz, _u64p, _u64p]),
    ("whir_gpu_sumcheck_fold", _int, [_u64p, _int, _sz, _u64p, _u64p]),
    ("whir_dev_rs_encode", _int, [_vp, _vp, _sz, _sz, _sz, _sz, _vp]),
    ("whir_dev_rs_encode_ext3", _int, [_vp, _vp, _sz, _sz, _sz, _sz, _vp]),
    ("whir_dev_ntt_batch", _int, [_vp, _vp, _sz, _sz, _int]),
    ("whir_dev_ntt_batch_transpose", _int, [_vp, _vp, _vp, _sz, _sz, _sz, _sz]),
    ("whir_dev_commit_leaves", _int, [_vp, _vp, _sz, _sz, _vp]),
    ("whir_dev_blake3_hash_many", _int, [_vp, _sz, _vp, _sz, _vp]),
    ("whir_merkle_num_nodes", _sz, [_sz]),
    ("whir_dev_merkle_build", _int, [_vp, _vp, _sz]),
    ("whir_dev_gather_hashes", _int, [_vp, _vp, _vp, _sz, _vp]),
    ("whir_dev_commit_base", _int, [_vp, _vp, _sz, _sz, _sz, _sz, _vp, _vp, _u8p]),
    ("whir_dev_commit_ext3", _int, [_vp, _vp, _sz, _sz, _sz, _sz, _vp, _vp, _u8p]),
    ("whir_dev_sumcheck_compute", _int, [_vp, _vp, _int, _vp, _sz, _u64p, _u64p]),
    ("whir_dev_sumcheck_fold", _int, [_vp, _vp, _int, _sz, _u64p, _vp]),
    ("whir_dev_horner", _int, [_vp, _vp, _int, _sz, _u64p, _u64p]),
    ("whir_dev_sumcheck_round", _int, [_vp, _vp, _int, _vp, _sz, _u64p, _vp, _vp, _u64p, _u64p]),
    ("whir_dev_sumcheck_round_noalloc", _int, [_vp, _vp, _int, _vp, _sz, _u64p, _vp, _vp, _vp]),
    ("whir_params_default", None, [_vp]),
    ("whir_config_flat", _sz, [ctypes.c_uint64, _vp, _u64p, _sz]),
    ("whir_domain_separator_cli", _int, [_vp, ctypes.c_char_p, _u8p, _u8p]),
    ("whir_prover_create", _int, [_vp, _vp, ctypes.c_uint64, _u8p, _u8p, _u8p, _sz, ctypes.POINTER(_vp)]),
    ("whir_prover_create_in", _int, [_vp, _vp, ctypes.c_uint64, _u8p, _u8p, _u8p, _sz, _vp, _sz,
                                     ctypes.POINTER(_vp)]),
    ("whir_prover_workspace_bytes", _sz, [_vp, ctypes.c_uint64]),
    ("whir_prover_destroy", _int, [_vp]),
    ("whir_prover_commit", _int, [_vp, _vp, _sz]),
    ("whir_prover_prove", _int, [_vp, _sz, _u64p, _u64p, _u64p, _u64p]),
    ("whir_prover_proof", _int, [_vp, _u8p, _sz, ctypes.POINTER(_sz), _u8p, _sz, ctypes.POINTER(_sz)]),
    ("whir_prover_pow_hashes", _sz, [_vp]),
    ("whir_prover_phase_times", _int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    ("whir_transcript_create", _int, [_u8p, _u8p, _u8p, _sz, ctypes.POINTER(_vp)]),
    ("whir_transcript_destroy", _int, [_vp]),
    ("whir_transcript_absorb", _int, [_vp, _u8p, _sz]),
    ("whir_transcript_message", _int, [_vp, _u8p, _sz]),
    ("whir_transcript_hint", _int, [_vp, _u8p, _sz]),
    ("whir_transcript_hint_vector", _int, [_vp, _u8p, _sz, _sz]),
    ("whir_transcript_squeeze", _int, [_vp, _u8p, _sz]),
    ("whir_transcript_challenge_gl", _int, [_vp, _u64p]),
    ("whir_transcript_indices", _sz, [_vp, _sz, _sz, _int, _u64p]),
    ("whir_transcript_proof", _int, [_vp, _u8p, _sz, ctypes.POINTER(_sz), _u8p, _sz, ctypes.POINTER(_sz)]),
    ("whir_dev_pow_search", _int, [_vp, _u8p, ctypes.c_uint64, _u64p]),
    ("whir_dev_eq_accumulate", _int, [_vp, _vp, _sz, _u64p, _u64p, _int]),
    ("whir_dev_geometric_accumulate", _int, [_vp, _vp, _sz, _u64p, _u64p, _sz]),
    ("whir_dev_mle_evaluate", _int, [_vp, _vp, _int, _sz, _u64p, _u64p]),
    ("whir_dev_mle_evaluate_many", _int, [_vp, _vp, _int, _sz, _u64p, _sz, _u64p]),
    ("whir_autotune_lde", _int, [_vp, _sz, _sz, _sz, _int, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_uint32)]),
    ("whir_tuning_export", _sz, [_u64p, _sz]),
    ("whir_tuning_import", _int, [_u64p, _sz]),
    ("whir_batch_create", _int, [_vp, _sz, _u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _int,
                                 ctypes.POINTER(_vp)]),
    ("whir_batch_create_form", _int, [_vp, _sz, _u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _int,
                                      _int, ctypes.POINTER(_vp)]),
    ("whir_batch_form", _int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("whir_batch_destroy", _int, [_vp]),
    ("whir_batch_coeffs", _vp, [_vp, _sz]),
    ("whir_batch_matrix", _vp, [_vp, _sz]),
    ("whir_batch_nodes", _vp, [_vp, _sz]),
    ("whir_batch_rows", _sz, [_vp, _sz]),
    ("whir_batch_device_bytes", _sz, [_vp]),
    ("whir_batch_run", _int, [_vp]),
    ("whir_batch_run_eager", _int, [_vp]),
    ("whir_batch_roots", _int, [_vp, _u8p]),
    ("whir_server_create", _int, [_int, _vp, ctypes.c_uint32, _sz, ctypes.POINTER(_vp)]),
    ("whir_server_destroy", _int, [_vp]),
    ("whir_server_request_bytes", _sz, [_vp, ctypes.c_uint64]),
    ("whir_server_submit", _int, [_vp, ctypes.c_uint64, _u64p, _int, _sz, _u64p, _u64p, _u64p]),
    ("whir_server_inject_failure", _int, [_vp, ctypes.c_uint64, ctypes.c_uint32]),
    ("whir_host_malloc", _int, [_sz, ctypes.POINTER(_vp)]),
    ("whir_host_free", _int, [_vp]),
    ("whir_server_wait", _int, [_vp, ctypes.c_uint64, _u8p, _sz, ctypes.POINTER(_sz), _u8p, _sz, ctypes.POINTER(_sz),
                                ctypes.POINTER(ctypes.c_double)]),
    ("whir_server_stats", _int, [_vp, _u64p]),
    ("whir_ctx_enable_timing", _int, [_vp, _int]),
    ("whir_ctx_reset_timing", _int, [_vp]),
    ("whir_ctx_stage_times", _int, [_vp, ctypes.POINTER(ctypes.c_double), _u64p]),
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libwhir_gpu.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; build it with `python -m whirplonk_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def symbols() -> list[str]:
    return [s[0] for s in _SIGS]


def _check(rc: int) -> None:
    if rc != 0:
        raise WhirGpuError(rc, load().whir_last_error().decode())


def _p64(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def _p8(a: np.ndarray):
    return a.ctypes.data_as(_u8p)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def device_available() -> bool:
    return bool(load().whir_device_available())


# ------------------------------------------------------- host-pointer forms

def ntt_batch(values, size: int, inverse: bool = False) -> np.ndarray:
    v = _u64(values).copy()
    fn = load().whir_gpu_intt_batch if inverse else load().whir_gpu_ntt_batch
    _check(fn(_p64(v), v.size, size))
    return v


def ntt_batch_transpose(values, ntt_size: int, rows: int, cols: int) -> np.ndarray:
    """gpu_ntt_batch_transpose: NTT of every ntt_size block, then the rows x
    cols result transposed (cols x rows, row-major)."""
    v = _u64(values).copy()
    _check(load().whir_gpu_ntt_batch_transpose(_p64(v), v.size, ntt_size, rows, cols))
    return v


def interleaved_rs_encode(coeffs, num_polys, poly_size, codeword_length, depth) -> np.ndarray:
    c = _u64(coeffs)
    out = np.empty(codeword_length * num_polys * depth, dtype=np.uint64)
    _check(load().whir_gpu_interleaved_rs_encode(_p64(c), _p64(out), num_polys, poly_size, codeword_length, depth))
    return out


def rs_encode_leaves(coeffs, num_polys, poly_size, codeword_length, depth):
    c = _u64(coeffs)
    out = np.empty(codeword_length * num_polys * depth, dtype=np.uint64)
    leaves = np.empty(32 * codeword_length, dtype=np.uint8)
    _check(load().whir_gpu_interleaved_rs_encode_blake3_matrix_leaves(
        _p64(c), _p64(out), _p8(leaves), num_polys, poly_size, codeword_length, depth))
    return out, leaves


def rs_encode_leaves_ext3(coeffs_aos, num_polys, poly_size, codeword_length, depth):
    """Ext3 AoS coefficients in; AoS matrix and leaves out (the ABI is SoA)."""
    c = _u64(coeffs_aos)
    comps = [np.ascontiguousarray(c[k::3]) for k in range(3)]
    n = codeword_length * num_polys * depth
    outs = [np.empty(n, dtype=np.uint64) for _ in range(3)]
    leaves = np.empty(32 * codeword_length, dtype=np.uint8)
    _check(load().whir_gpu_interleaved_rs_encode_blake3_ext3_matrix_leaves(
        _p64(comps[0]), _p64(comps[1]), _p64(comps[2]), _p64(outs[0]), _p64(outs[1]), _p64(outs[2]),
        _p8(leaves), num_polys, poly_size, codeword_length, depth))
    aos = np.empty(3 * n, dtype=np.uint64)
    for k in range(3):
        aos[k::3] = outs[k]
    return aos, leaves


HASH_BLAKE3 = 1
HASH_SHA2 = 2


def merkle_tree(leaves: np.ndarray, num_leaves: int, full: bool = True, hash: int = HASH_BLAKE3):
    lv = np.ascontiguousarray(leaves, dtype=np.uint8)
    if lv.size == 0:
        lv = np.zeros(32, dtype=np.uint8)
    root = np.empty(32, dtype=np.uint8)
    nodes = np.empty(32 * load().whir_merkle_num_nodes(num_leaves), dtype=np.uint8) if full else None
    fn = load().whir_gpu_sha256_merkle_tree if hash == HASH_SHA2 else load().whir_gpu_blake3_merkle_tree
    _check(fn(_p8(lv), num_leaves, _p8(root), _p8(nodes) if full else None))
    return root, nodes


def merkle_hint_node_indices(num_leaves: int, indices) -> np.ndarray:
    ix = _u64(indices) if len(indices) else np.zeros(1, dtype=np.uint64)
    n = load().whir_merkle_hint_node_indices(num_leaves, _p64(ix), len(indices), None)
    out = np.empty(max(n, 1), dtype=np.uint64)
    load().whir_merkle_hint_node_indices(num_leaves, _p64(ix), len(indices), _p64(out))
    return out[:n]


def merkle_open_path(leaves: np.ndarray, num_leaves: int, indices, hash: int = HASH_BLAKE3):
    lv = np.ascontiguousarray(leaves, dtype=np.uint8)
    ix = _u64(indices) if len(indices) else np.zeros(1, dtype=np.uint64)
    nh = len(merkle_hint_node_indices(num_leaves, indices))
    root = np.empty(32, dtype=np.uint8)
    hints = np.empty(max(32 * nh, 32), dtype=np.uint8)
    fn = load().whir_gpu_sha256_merkle_open_path if hash == HASH_SHA2 else load().whir_gpu_blake3_merkle_open_path
    _check(fn(_p8(lv), num_leaves, _p64(ix), len(indices), _p8(root), _p8(hints)))
    return root, hints[: 32 * nh]


def ood_evaluate_from_base(coeffs, points, num_vectors, vector_size) -> np.ndarray:
    c, pts = _u64(coeffs), _u64(points)
    npnt = pts.size // 3
    out = np.empty(3 * npnt * num_vectors, dtype=np.uint64)
    _check(load().whir_gpu_ood_evaluate_ext3_from_base(_p64(c), _p64(pts), _p64(out), num_vectors, vector_size, npnt))
    return out


def ood_evaluate_from_ext3(coeffs_aos, points, num_vectors, vector_size) -> np.ndarray:
    c, pts = _u64(coeffs_aos), _u64(points)
    comps = [np.ascontiguousarray(c[k::3]) for k in range(3)]
    npnt = pts.size // 3
    out = np.empty(3 * npnt * num_vectors, dtype=np.uint64)
    _check(load().whir_gpu_ood_evaluate_ext3_from_ext3(_p64(comps[0]), _p64(comps[1]), _p64(comps[2]), _p64(pts),
                                                       _p64(out), num_vectors, vector_size, npnt))
    return out


def blake3_hash_many(size: int, data: bytes, count: int) -> np.ndarray:
    buf = np.frombuffer(bytes(data) + b"\0" * 16, dtype=np.uint8)
    out = np.empty(max(32 * count, 32), dtype=np.uint8)
    _check(load().whir_gpu_blake3_hash_many(size, _p8(buf), count, _p8(out)))
    return out[: 32 * count]


def sha256_hash_many(size: int, data: bytes, count: int) -> np.ndarray:
    buf = np.frombuffer(bytes(data) + b"\0" * 16, dtype=np.uint8)
    out = np.empty(max(32 * count, 32), dtype=np.uint8)
    _check(load().whir_gpu_sha256_hash_many(size, _p8(buf), count, _p8(out)))
    return out[: 32 * count]


def sumcheck_compute(a, b, a_is_base: bool = False):
    a, b = _u64(a), _u64(b)
    n = b.size // 3
    c0, c2 = np.empty(3, dtype=np.uint64), np.empty(3, dtype=np.uint64)
    _check(load().whir_gpu_sumcheck_compute(_p64(a), int(a_is_base), _p64(b), n, _p64(c0), _p64(c2)))
    return c0, c2


def sumcheck_fold(v, r, v_is_base: bool = False) -> np.ndarray:
    v, r = _u64(v), _u64(r)
    n = v.size if v_is_base else v.size // 3
    out = np.empty(3 * (n // 2) + 1, dtype=np.uint64)
    _check(load().whir_gpu_sumcheck_fold(_p64(v), int(v_is_base), n, _p64(r), _p64(out)))
    return out[: 3 * (n // 2)]


# ------------------------------------------------------ device-resident forms

class Context:
    """An explicit device context (stream + pools) with device buffers."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(load().whir_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._bufs = []

    def close(self):
        if self.h:
            for b in self._bufs:
                load().whir_dev_free(self.h, b)
            self._bufs = []
            load().whir_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes: int) -> int:
        p = _vp()
        _check(load().whir_dev_malloc(self.h, nbytes, ctypes.byref(p)))
        self._bufs.append(p.value)
        return p.value

    def free(self, ptr: int) -> None:
        if ptr in self._bufs:
            self._bufs.remove(ptr)
        _check(load().whir_dev_free(self.h, ptr))

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(max(arr.nbytes, 1))
        _check(load().whir_memcpy_h2d(self.h, p, arr.ctypes.data, arr.nbytes))
        return p

    def download(self, ptr: int, count: int, dtype=np.uint64) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(load().whir_memcpy_d2h(self.h, out.ctypes.data, ptr, out.nbytes))
        return out

    def copy(self, dst: int, src: int, nbytes: int):
        _check(load().whir_dev_copy(self.h, dst, src, nbytes))

    def set_hash(self, hash: int):
        """Hash engine of this context's trees, row leaves and PoW (HASH_BLAKE3 or HASH_SHA2)."""
        _check(load().whir_ctx_set_hash(self.h, hash))

    def sync(self):
        _check(load().whir_ctx_synchronize(self.h))

    def rs_encode(self, d_coeffs, num_polys, poly_size, cw, depth, d_out, ext3=False):
        fn = load().whir_dev_rs_encode_ext3 if ext3 else load().whir_dev_rs_encode
        _check(fn(self.h, d_coeffs, num_polys, poly_size, cw, depth, d_out))

    def ntt_batch(self, d_values, total, size, inverse=False):
        _check(load().whir_dev_ntt_batch(self.h, d_values, total, size, int(inverse)))

    def ntt_batch_transpose(self, d_values, d_out, total, ntt_size, rows, cols):
        _check(load().whir_dev_ntt_batch_transpose(self.h, d_values, d_out, total, ntt_size, rows, cols))

    def commit(self, d_coeffs, num_polys, poly_size, cw, depth, d_matrix, d_nodes, ext3=False) -> bytes:
        root = np.empty(32, dtype=np.uint8)
        fn = load().whir_dev_commit_ext3 if ext3 else load().whir_dev_commit_base
        _check(fn(self.h, d_coeffs, num_polys, poly_size, cw, depth, d_matrix, d_nodes, _p8(root)))
        return root.tobytes()

    def commit_leaves(self, d_matrix, rows, row_u64, d_leaves):
        _check(load().whir_dev_commit_leaves(self.h, d_matrix, rows, row_u64, d_leaves))

    def merkle_build(self, d_nodes, num_leaves):
        _check(load().whir_dev_merkle_build(self.h, d_nodes, num_leaves))

    def gather_hashes(self, d_nodes, d_idx, count, d_out):
        _check(load().whir_dev_gather_hashes(self.h, d_nodes, d_idx, count, d_out))

    def sumcheck_compute(self, d_a, a_is_base, d_b, n):
        c0, c2 = np.empty(3, dtype=np.uint64), np.empty(3, dtype=np.uint64)
        _check(load().whir_dev_sumcheck_compute(self.h, d_a, int(a_is_base), d_b, n, _p64(c0), _p64(c2)))
        return c0, c2

    def sumcheck_fold(self, d_v, v_is_base, n, r, d_out):
        rr = _u64(r)
        _check(load().whir_dev_sumcheck_fold(self.h, d_v, int(v_is_base), n, _p64(rr), d_out))

    def sumcheck_round(self, d_a, a_is_base, d_b, n, r, d_a_out, d_b_out):
        rr = _u64(r)
        c0, c2 = np.empty(3, dtype=np.uint64), np.empty(3, dtype=np.uint64)
        _check(load().whir_dev_sumcheck_round(self.h, d_a, int(a_is_base), d_b, n, _p64(rr), d_a_out, d_b_out,
                                              _p64(c0), _p64(c2)))
        return c0, c2

    def enable_timing(self, on: bool = True):
        _check(load().whir_ctx_enable_timing(self.h, int(on)))

    def reset_timing(self):
        _check(load().whir_ctx_reset_timing(self.h))

    STAGES = ("rs_encode", "leaves", "merkle", "sumcheck_round")

    def stage_times(self) -> dict:
        ms = (ctypes.c_double * 4)()
        cnt = np.zeros(4, dtype=np.uint64)
        _check(load().whir_ctx_stage_times(self.h, ms, _p64(cnt)))
        return {k: (ms[i], int(cnt[i])) for i, k in enumerate(self.STAGES)}

    def horner(self, d_f, f_is_ext3, n, z) -> np.ndarray:
        zz = _u64(z)
        out = np.empty(3, dtype=np.uint64)
        _check(load().whir_dev_horner(self.h, d_f, int(f_is_ext3), n, _p64(zz), _p64(out)))
        return out


    # ---- prove-path building blocks
    def pow_search(self, challenge: bytes, threshold: int) -> int:
        ch = np.frombuffer(bytes(challenge), dtype=np.uint8).copy()
        out = np.zeros(1, dtype=np.uint64)
        _check(load().whir_dev_pow_search(self.h, _p8(ch), threshold, _p64(out)))
        return int(out[0])

    def eq_accumulate(self, d_acc, point, scalar=(1, 0, 0), accumulate=True):
        pt = _u64(point).reshape(-1)
        sc = _u64(scalar)
        _check(load().whir_dev_eq_accumulate(self.h, d_acc, pt.size // 3, _p64(pt), _p64(sc), int(accumulate)))

    def geometric_accumulate(self, d_acc, n, points, scalars):
        pts, scs = _u64(points).reshape(-1), _u64(scalars).reshape(-1)
        _check(load().whir_dev_geometric_accumulate(self.h, d_acc, n, _p64(pts), _p64(scs), pts.size // 3))

    def mle_evaluate(self, d_f, f_is_base, point) -> np.ndarray:
        pt = _u64(point).reshape(-1)
        out = np.empty(3, dtype=np.uint64)
        _check(load().whir_dev_mle_evaluate(self.h, d_f, int(f_is_base), pt.size // 3, _p64(pt), _p64(out)))
        return out

    def mle_evaluate_many(self, d_f, f_is_base, points) -> np.ndarray:
        """The evaluations at points (count x num_variables x 3) -> (count, 3)."""
        pts = _u64(points)
        if pts.ndim != 3 or pts.shape[2] != 3:
            raise ValueError("points must have shape (count, num_variables, 3)")
        count, nv = pts.shape[0], pts.shape[1]
        out = np.empty((count, 3), dtype=np.uint64)
        flat = np.ascontiguousarray(pts.reshape(-1)) if pts.size else np.zeros(1, dtype=np.uint64)
        _check(load().whir_dev_mle_evaluate_many(self.h, d_f, int(f_is_base), nv, _p64(flat), count,
                                                 _p64(out) if count else None))
        return out


# ------------------------------------------------------------ prove loop

class Params(ctypes.Structure):
    """whir_params (ProtocolParameters, parameters.hpp:25-34)."""
    _fields_ = [
        ("unique_decoding", ctypes.c_int),
        ("starting_log_inv_rate", ctypes.c_uint32),
        ("initial_folding_factor", ctypes.c_uint32),
        ("folding_factor", ctypes.c_uint32),
        ("security_level", ctypes.c_uint32),
        ("pow_bits", ctypes.c_uint32),
        ("batch_size", ctypes.c_uint32),
        ("hash_id", ctypes.c_uint32),
    ]

    @classmethod
    def default(cls, **kw) -> "Params":
        p = cls()
        load().whir_params_default(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def workspace_bytes(size: int, params: Params) -> int:
    return load().whir_prover_workspace_bytes(ctypes.byref(params), size)


def config_flat(size: int, params: Params) -> np.ndarray:
    """whir::Config::from_params flattened (layout in include/whir_gpu.h)."""
    n = load().whir_config_flat(size, ctypes.byref(params), None, 0)
    if n == 0:
        raise WhirGpuError(3, "invalid parameters")
    out = np.zeros(n, dtype=np.uint64)
    load().whir_config_flat(size, ctypes.byref(params), _p64(out), n)
    return out


def domain_separator_cli(params: Params, session: str | None = None) -> tuple[bytes, bytes]:
    pid = np.zeros(64, dtype=np.uint8)
    sid = np.zeros(32, dtype=np.uint8)
    _check(load().whir_domain_separator_cli(ctypes.byref(params), session.encode() if session else None,
                                            _p8(pid), _p8(sid)))
    return pid.tobytes(), sid.tobytes()


def _buf(b: bytes) -> np.ndarray:
    return np.frombuffer(bytes(b), dtype=np.uint8).copy() if b else np.zeros(1, dtype=np.uint8)


class Transcript:
    """The host prover transcript (whir_transcript_*), for host-side checks."""

    def __init__(self, protocol_id: bytes | None = None, session_id: bytes | None = None, instance: bytes = b""):
        h = _vp()
        pid = _buf(protocol_id) if protocol_id is not None else None
        sid = _buf(session_id) if session_id is not None else None
        ins = _buf(instance)
        _check(load().whir_transcript_create(_p8(pid) if pid is not None else None,
                                             _p8(sid) if sid is not None else None, _p8(ins), len(instance),
                                             ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            load().whir_transcript_destroy(self.h)
            self.h = None

    def absorb(self, data: bytes):
        b = _buf(data)
        _check(load().whir_transcript_absorb(self.h, _p8(b), len(data)))

    def message(self, data: bytes):
        b = _buf(data)
        _check(load().whir_transcript_message(self.h, _p8(b), len(data)))

    def hint(self, data: bytes):
        b = _buf(data)
        _check(load().whir_transcript_hint(self.h, _p8(b), len(data)))

    def hint_vector(self, data: bytes, count: int, elem_bytes: int):
        b = _buf(data)
        _check(load().whir_transcript_hint_vector(self.h, _p8(b), count, elem_bytes))

    def squeeze(self, n: int) -> bytes:
        out = np.zeros(max(n, 1), dtype=np.uint8)
        _check(load().whir_transcript_squeeze(self.h, _p8(out), n))
        return out[:n].tobytes()

    def challenge_gl(self) -> int:
        out = np.zeros(1, dtype=np.uint64)
        _check(load().whir_transcript_challenge_gl(self.h, _p64(out)))
        return int(out[0])

    def indices(self, num_leaves: int, count: int, dedup: bool) -> list[int]:
        out = np.zeros(max(count, 1), dtype=np.uint64)
        k = load().whir_transcript_indices(self.h, num_leaves, count, int(dedup), _p64(out))
        return [int(x) for x in out[:k]]

    def proof(self) -> tuple[bytes, bytes]:
        na, nh = _sz(), _sz()
        _check(load().whir_transcript_proof(self.h, None, 0, ctypes.byref(na), None, 0, ctypes.byref(nh)))
        a = np.zeros(max(na.value, 1), dtype=np.uint8)
        b = np.zeros(max(nh.value, 1), dtype=np.uint8)
        _check(load().whir_transcript_proof(self.h, _p8(a), a.size, ctypes.byref(na), _p8(b), b.size,
                                            ctypes.byref(nh)))
        return a[:na.value].tobytes(), b[:nh.value].tobytes()


class Prover:
    """One proof: whir_prover_create / commit / prove / proof."""

    def __init__(self, ctx: Context, params: Params, size: int, protocol_id: bytes, session_id: bytes,
                 instance: bytes = b"", workspace: tuple[int, int] | None = None):
        """workspace=(device pointer, bytes) carves the proof's memory from a
        caller-owned region (whir_prover_create_in) instead of allocating."""
        self.ctx = ctx
        self.size = size
        pid = np.frombuffer(protocol_id, dtype=np.uint8).copy()
        sid = np.frombuffer(session_id, dtype=np.uint8).copy()
        ins = np.frombuffer(instance, dtype=np.uint8).copy() if instance else np.zeros(1, dtype=np.uint8)
        h = _vp()
        if workspace is None:
            _check(load().whir_prover_create(ctx.h, ctypes.byref(params), size, _p8(pid), _p8(sid), _p8(ins),
                                             len(instance), ctypes.byref(h)))
        else:
            _check(load().whir_prover_create_in(ctx.h, ctypes.byref(params), size, _p8(pid), _p8(sid), _p8(ins),
                                                len(instance), workspace[0], workspace[1], ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().whir_prover_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def commit(self, d_vectors: int, num_vectors: int = 1):
        _check(load().whir_prover_commit(self.h, d_vectors, num_vectors))

    def prove(self, points, evaluations):
        """points: (num_forms, num_variables, 3); evaluations: (num_forms * num_vectors, 3)."""
        pts = _u64(points).reshape(-1)
        ev = _u64(evaluations).reshape(-1)
        nv = self.size.bit_length() - 1
        nf = pts.size // (3 * nv) if nv else 0
        ep = np.zeros(3 * nv, dtype=np.uint64)
        rl = np.zeros(3 * max(nf, 1), dtype=np.uint64)
        _check(load().whir_prover_prove(self.h, nf, _p64(pts) if nf else None, _p64(ev) if nf else None,
                                        _p64(ep), _p64(rl)))
        return ep.reshape(-1, 3), rl[:3 * nf].reshape(-1, 3)

    def proof(self) -> tuple[bytes, bytes]:
        na, nh = _sz(), _sz()
        _check(load().whir_prover_proof(self.h, None, 0, ctypes.byref(na), None, 0, ctypes.byref(nh)))
        a = np.zeros(max(na.value, 1), dtype=np.uint8)
        b = np.zeros(max(nh.value, 1), dtype=np.uint8)
        _check(load().whir_prover_proof(self.h, _p8(a), a.size, ctypes.byref(na), _p8(b), b.size, ctypes.byref(nh)))
        return a[:na.value].tobytes(), b[:nh.value].tobytes()

    def pow_hashes(self) -> int:
        return load().whir_prover_pow_hashes(self.h)

    PHASES = ("commit", "ood", "weights", "sumcheck", "pow", "round_commit", "open", "other")

    def phase_times(self) -> dict:
        ms = (ctypes.c_double * len(self.PHASES))()
        _check(load().whir_prover_phase_times(self.h, ms))
        return {k: ms[i] for i, k in enumerate(self.PHASES)}


def autotune_lde(ctx: "Context", m: int, n: int, columns: int = 16, ext3: bool = False) -> tuple[float, int]:
    ms = ctypes.c_double()
    lg = ctypes.c_uint32()
    _check(load().whir_autotune_lde(ctx.h, m, n, columns, int(ext3), ctypes.byref(ms), ctypes.byref(lg)))
    return ms.value, lg.value


def tuning_export() -> np.ndarray:
    n = load().whir_tuning_export(None, 0)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    load().whir_tuning_export(_p64(out), n)
    return out[:n]


def tuning_import(records) -> None:
    r = _u64(records)
    _check(load().whir_tuning_import(_p64(r), r.size))


class Batch:
    """Batched commits of many polynomials in one graph launch (whir_batch_*)."""

    def __init__(self, ctx: "Context", sizes, log_inv_rate: int = 2, folding_factor: int = 4,
                 branches: int = 8, autotune: bool = True, form: int = -1):
        """form: 1 level-synchronous encodes, 0 one commit per polynomial,
        -1 the library's choice (timed when autotune is set)."""
        self.ctx = ctx
        self.sizes = [int(s) for s in sizes]
        sz = _u64(self.sizes)
        h = _vp()
        _check(load().whir_batch_create_form(ctx.h, len(self.sizes), _p64(sz), log_inv_rate, folding_factor,
                                             branches, int(autotune), int(form), ctypes.byref(h)))
        self.h = h

    def form(self):
        """(form, per-polynomial ms, level-synchronous ms): the form the batch
        runs (1 level-synchronous, 0 per polynomial) and, when create timed
        both, their graph replay times (else -1)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        f = load().whir_batch_form(self.h, ctypes.byref(a), ctypes.byref(b))
        return f, a.value, b.value

    def close(self):
        if self.h:
            load().whir_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, k: int, coeffs):
        c = _u64(coeffs)
        _check(load().whir_memcpy_h2d(self.ctx.h, load().whir_batch_coeffs(self.h, k), c.ctypes.data, c.nbytes))

    def matrix(self, k: int) -> np.ndarray:
        rows = load().whir_batch_rows(self.h, k)
        return self.ctx.download(load().whir_batch_matrix(self.h, k), rows * 16)

    def nodes(self, k: int) -> np.ndarray:
        """Polynomial k's flat Merkle node array (32 bytes per node)."""
        rows = load().whir_batch_rows(self.h, k)
        return self.ctx.download(load().whir_batch_nodes(self.h, k), 4 * merkle_num_nodes(rows)).view(np.uint8)

    def run(self, graph: bool = True):
        _check((load().whir_batch_run if graph else load().whir_batch_run_eager)(self.h))

    def roots(self) -> list[bytes]:
        out = np.zeros(32 * len(self.sizes), dtype=np.uint8)
        _check(load().whir_batch_roots(self.h, _p8(out)))
        return [out[32 * k:32 * k + 32].tobytes() for k in range(len(self.sizes))]

    def device_bytes(self) -> int:
        return load().whir_batch_device_bytes(self.h)


class Pinned:
    """Pinned host memory (whir_host_malloc) viewed as a numpy array; freed
    when this object is collected, so keep it alive while the array is used."""

    def __init__(self, count: int, dtype=np.uint64):
        nbytes = max(int(count) * np.dtype(dtype).itemsize, 1)
        p = _vp()
        _check(load().whir_host_malloc(nbytes, ctypes.byref(p)))
        self.ptr = p.value
        self.array = np.frombuffer((ctypes.c_uint8 * nbytes).from_address(self.ptr), dtype=dtype, count=count)

    def __del__(self):
        if getattr(self, "ptr", None):
            load().whir_host_free(self.ptr)
            self.ptr = None


class Server:
    """The serving loop (whir_server_*): submit requests, wait for proofs."""

    def __init__(self, params: Params, num_slots: int = 4, hbm_budget: int = 8 << 30, device: int = 0):
        h = _vp()
        _check(load().whir_server_create(device, ctypes.byref(params), num_slots, hbm_budget, ctypes.byref(h)))
        self.h = h
        # request id -> the array a borrowed request reads in place; held here
        # until its wait() returns, since it may be a converted temporary
        self._borrowed: dict[int, np.ndarray] = {}

    def close(self):
        if self.h:
            _check(load().whir_server_destroy(self.h))  # finishes every request first
            self.h = None
            self._borrowed.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def request_bytes(self, size: int) -> int:
        return load().whir_server_request_bytes(self.h, size)

    def submit(self, size: int, coeffs, points=None, evaluations=None, borrow: bool = False) -> int:
        """borrow=True lets the server read the coefficients in place, without
        a copy at submit.  The array it reads (coeffs itself when it is already
        a C-contiguous uint64 array, else the converted copy) is kept alive by
        this object until the request's wait() returns."""
        c = _u64(coeffs)
        pts = _u64(points).reshape(-1) if points is not None and np.size(points) else None
        nf = pts.size // (3 * (size.bit_length() - 1)) if pts is not None else 0
        ev = _u64(evaluations).reshape(-1) if evaluations is not None else None
        out = np.zeros(1, dtype=np.uint64)
        _check(load().whir_server_submit(self.h, size, _p64(c), int(borrow), nf, _p64(pts) if nf else None,
                                         _p64(ev) if ev is not None else None, _p64(out)))
        rid = int(out[0])
        if borrow:
            self._borrowed[rid] = c
        return rid

    def inject_failure(self, rid: int, at_wait: int) -> None:
        """Test hook: request rid (submitted or not yet) fails at its
        at_wait-th device wait, with work still in flight (0 clears)."""
        _check(load().whir_server_inject_failure(self.h, rid, at_wait))

    def wait(self, rid: int):
        """(narg, hints, (queue_ms, service_ms, latency_ms))"""
        na, nh = _sz(), _sz()
        ms = (ctypes.c_double * 3)()
        try:
            _check(load().whir_server_wait(self.h, rid, None, 0, ctypes.byref(na), None, 0, ctypes.byref(nh), ms))
        except WhirGpuError:
            self._borrowed.pop(rid, None)  # the request failed and was released
            raise
        self._borrowed.pop(rid, None)  # done: the server no longer reads the input
        a = np.zeros(max(na.value, 1), dtype=np.uint8)
        b = np.zeros(max(nh.value, 1), dtype=np.uint8)
        _check(load().whir_server_wait(self.h, rid, _p8(a), a.size, ctypes.byref(na), _p8(b), b.size,
                                       ctypes.byref(nh), ms))
        return a[:na.value].tobytes(), b[:nh.value].tobytes(), (ms[0], ms[1], ms[2])

    STATS = ("completed", "running", "queued", "in_use", "peak", "budget", "slots", "pow_launches", "pow_searches",
             "encode_launches", "encodes")

    def stats(self) -> dict:
        out = np.zeros(len(self.STATS), dtype=np.uint64)
        _check(load().whir_server_stats(self.h, _p64(out)))
        return {k: int(v) for k, v in zip(self.STATS, out)}


def merkle_num_nodes(n: int) -> int:
    return load().whir_merkle_num_nodes(n)


# ------------------------------------------------------------ seeded inputs

_LCG_A = 6364136223846793005
_LCG_C = 1442695040888963407


def lcg_stream(seed: int, n: int) -> np.ndarray:
    """The first n states of the LCG x <- x*A + C (mod 2^64) used by the
    reference's tests and dumpers (e.g. cpp/tests/dump_whir.cpp:43-50),
    vectorised with jump-ahead: block [2^j, 2^(j+1)) = A^(2^j) * block[0, 2^j) + C_(2^j)."""
    out = np.empty(max(n, 1), dtype=np.uint64)
    with np.errstate(over="ignore"):
        out[0] = np.uint64((seed * _LCG_A + _LCG_C) % 2**64)
        have, a, c = 1, _LCG_A, _LCG_C  # (a, c) maps x_k -> x_{k+have}
        while have < n:
            take = min(have, n - have)
            out[have:have + take] = out[:take] * np.uint64(a) + np.uint64(c)
            a, c = (a * a) % 2**64, (a * c + c) % 2**64
            have += take
    return out[:n]


def lcg_uniform(seed: int, n: int) -> tuple[np.ndarray, int]:
    """n values uniform on [0, p): the LCG stream with draws >= p rejected.
    Returns (values, state) so streams can be chained like the reference's Lcg."""
    want = n + 64
    while True:
        s = lcg_stream(seed, want)
        keep = s[s < np.uint64(P)]
        if keep.size >= n:
            # the state after the n-th accepted draw
            last = int(np.nonzero(s < np.uint64(P))[0][n - 1]) if n else -1
            state = int(s[last]) if n else seed
            return keep[:n].copy(), state
        want += 1024
