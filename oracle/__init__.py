"""Python bindings for the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this package; the product (whirplonk_amd) never does.  It wraps

* ``oracle/_build/liboracle.so`` — the C restatement in whir_oracle.c and
  keccak_oracle.c (built by ``make -C oracle``), and
* ``oracle/_ref/libwhirref.so`` — the reference C++ prover's own functions,
  compiled from its sources by ``make -C oracle ref`` (optional; present only
  where the reference checkout existed at build time).

All field data crosses as numpy uint64 arrays (Ext3 as (n, 3) AoS) and hashes
as uint8 arrays of 32 bytes per hash.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libwhirref.so")

P = 0xFFFFFFFF00000001

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t
_u64 = ctypes.c_uint64


def build(ref: bool = False) -> None:
    """Compile the oracle (and, if asked and the reference exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/cpp/include"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
        subprocess.run(["make", "-s", "-C", HERE, "integration"], check=True)


def _ptr64(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def _ptr8(a: np.ndarray):
    return a.ctypes.data_as(_u8p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name in ("or_gl_from_u64", "or_gl_neg", "or_gl_inv", "or_gl_root"):
            getattr(L, name).restype = _u64
            getattr(L, name).argtypes = [_u64]
        for name in ("or_gl_add", "or_gl_sub", "or_gl_mul", "or_gl_pow"):
            getattr(L, name).restype = _u64
            getattr(L, name).argtypes = [_u64, _u64]
        for name in ("or_e3_add", "or_e3_sub", "or_e3_mul"):
            getattr(L, name).argtypes = [_u64p, _u64p, _u64p]
        L.or_e3_pow.argtypes = [_u64p, _u64, _u64p]
        L.or_e3_inv.argtypes = [_u64p, _u64p]
        L.or_ntt_batch.argtypes = [_u64p, _sz, _sz]
        L.or_intt_batch.argtypes = [_u64p, _sz, _sz]
        L.or_dft_naive.argtypes = [_u64p, _u64p, _sz]
        L.or_rs_encode.argtypes = [_u64p, _sz, _sz, _sz, _sz, _u64p]
        L.or_rs_encode_ext3.argtypes = [_u64p, _sz, _sz, _sz, _sz, _u64p]
        L.or_blake3.argtypes = [_u8p, _sz, _u8p]
        L.or_blake3_many.argtypes = [_sz, _u8p, _sz, _u8p]
        L.or_commit_leaves_base.argtypes = [_u64p, _sz, _sz, _u8p]
        L.or_commit_leaves_ext3.argtypes = [_u64p, _sz, _sz, _u8p]
        L.or_merkle_layers.restype = _sz
        L.or_merkle_layers.argtypes = [_sz]
        L.or_merkle_num_nodes.restype = _sz
        L.or_merkle_num_nodes.argtypes = [_sz]
        L.or_merkle_build.argtypes = [_u8p, _sz, _u8p]
        L.or_merkle_open.restype = _sz
        L.or_merkle_open.argtypes = [_u8p, _sz, _u64p, _sz, _u8p]
        L.or_merkle_verify.restype = ctypes.c_int
        L.or_merkle_verify.argtypes = [_u8p, _sz, _u64p, _u8p, _sz, _u8p, _sz]
        L.or_sumcheck_compute_ext3.argtypes = [_u64p, _u64p, _sz, _u64p, _u64p]
        L.or_sumcheck_compute_base_ext3.argtypes = [_u64p, _u64p, _sz, _u64p, _u64p]
        L.or_sumcheck_fold_ext3.argtypes = [_u64p, _sz, _u64p]
        L.or_sumcheck_fold_base_ext3.argtypes = [_u64p, _sz, _u64p, _u64p]
        L.or_horner_base_ext3.argtypes = [_u64p, _sz, _u64p, _u64p]
        L.or_horner_ext3_ext3.argtypes = [_u64p, _sz, _u64p, _u64p]
        L.or_horner_base.restype = _u64
        L.or_horner_base.argtypes = [_u64p, _sz, _u64]
        L.or_lcg_fill.argtypes = [_u64p, _u64p, _sz, ctypes.c_int]
        L.or_eval_eq_ext3.argtypes = [_u64p, _u64p, _sz, _u64p]
        L.or_geometric_accumulate_ext3.argtypes = [_u64p, _sz, _u64p, _u64p, _sz]
        L.or_mle_evaluate.argtypes = [_u64p, ctypes.c_int, _u64p, _sz, _u64p]
        L.or_pow_find.restype = _u64
        L.or_pow_find.argtypes = [_u8p, _u64]
        L.or_sha3_256.argtypes = [_u8p, _sz, _u8p]
        L.or_sha3_512.argtypes = [_u8p, _sz, _u8p]
        L.or_shake128.argtypes = [_u8p, _sz, _u8p, _sz]
        _lib = L
    return _lib


def ref():
    """The reference C++ build, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        R.ref_num_threads.restype = ctypes.c_int
        R.ref_ntt_batch.argtypes = [_u64p, _sz, _sz]
        R.ref_rs_encode.argtypes = [_u64p, _sz, _sz, _sz, _sz, _u64p]
        R.ref_rs_encode_ext3.argtypes = [_u64p, _sz, _sz, _sz, _sz, _u64p]
        R.ref_commit_leaves_base.argtypes = [_u64p, _sz, _sz, _u8p]
        R.ref_commit_leaves_ext3.argtypes = [_u64p, _sz, _sz, _u8p]
        R.ref_merkle_build.argtypes = [_u8p, _sz, _u8p]
        R.ref_merkle_open.restype = _sz
        R.ref_merkle_open.argtypes = [_u8p, _sz, _u64p, _sz, _u8p]
        R.ref_sumcheck_compute_ext3.argtypes = [_u64p, _u64p, _sz, _u64p, _u64p]
        R.ref_sumcheck_fold_ext3.argtypes = [_u64p, _sz, _u64p]
        R.ref_horner_base_ext3.argtypes = [_u64p, _sz, _u64p, _u64p]
        R.ref_horner_ext3_ext3.argtypes = [_u64p, _sz, _u64p, _u64p]
        R.ref_pow_find.restype = _u64
        R.ref_pow_find.argtypes = [ctypes.c_uint32, _u8p, _u64]
        R.ref_e3_pow.argtypes = [_u64p, _u64, _u64p]
        R.ref_commit_base.argtypes = [_u64p, _sz, _sz, _sz, _u64p, _u8p]
        R.ref_blake3_many.argtypes = [_sz, _u8p, _sz, _u8p]
        R.ref_whir_config.restype = _sz
        R.ref_whir_config.argtypes = [ctypes.c_uint32] * 5 + [_u64p, _sz]
        R.ref_whir_prove_params.restype = ctypes.c_int
        R.ref_whir_prove_params.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, _u64p, _sz, _u64p,
                                            _u64p, _u8p, _sz, ctypes.POINTER(_sz), _u8p, _sz, ctypes.POINTER(_sz)]
        R.ref_cli_ds.restype = None
        R.ref_cli_ds.argtypes = [ctypes.POINTER(ctypes.c_uint32), _u8p, _u8p]
        R.ref_whir_prove_hash.restype = ctypes.c_int
        R.ref_whir_prove_hash.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32, _u64p,
                                          _sz, _u64p, _u64p, _u8p, _sz, ctypes.POINTER(_sz), _u8p, _sz,
                                          ctypes.POINTER(_sz)]
        R.ref_cli_ds_hash.restype = None
        R.ref_cli_ds_hash.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, _u8p, _u8p]
        R.ref_whir_verify_hash.restype = ctypes.c_int
        R.ref_whir_verify_hash.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32, _sz,
                                           _u64p, _u64p, _u8p, _sz, _u8p, _sz]
        R.ref_sha256_many.argtypes = [_sz, _u8p, _sz, _u8p]
        R.ref_merkle_build_sha2.argtypes = [_u8p, _sz, _u8p]
        R.ref_commit_leaves_base_sha2.argtypes = [_u64p, _sz, _sz, _u8p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers

def u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def lcg(seed: int, n: int, mode: int = 0) -> tuple[np.ndarray, int]:
    """n draws of the reference LCG; mode 0 = from_u64, mode 1 = reject x >= p."""
    st = np.array([seed], dtype=np.uint64)
    out = np.empty(n, dtype=np.uint64)
    lib().or_lcg_fill(_ptr64(st), _ptr64(out), n, mode)
    return out, int(st[0])


def gl_add(a, b): return lib().or_gl_add(a, b)
def gl_sub(a, b): return lib().or_gl_sub(a, b)
def gl_mul(a, b): return lib().or_gl_mul(a, b)
def gl_pow(a, e): return lib().or_gl_pow(a, e)
def gl_inv(a): return lib().or_gl_inv(a)
def gl_root(n): return lib().or_gl_root(n)


def _e3op(fn, a, b):
    a, b, o = u64(a), u64(b), np.zeros(3, dtype=np.uint64)
    fn(_ptr64(a), _ptr64(b), _ptr64(o))
    return o


def e3_add(a, b): return _e3op(lib().or_e3_add, a, b)
def e3_sub(a, b): return _e3op(lib().or_e3_sub, a, b)
def e3_mul(a, b): return _e3op(lib().or_e3_mul, a, b)


def e3_pow(a, e):
    a, o = u64(a), np.zeros(3, dtype=np.uint64)
    lib().or_e3_pow(_ptr64(a), e, _ptr64(o))
    return o


def e3_inv(a):
    a, o = u64(a), np.zeros(3, dtype=np.uint64)
    lib().or_e3_inv(_ptr64(a), _ptr64(o))
    return o


def ntt_batch(values, n: int) -> np.ndarray:
    v = u64(values).copy()
    lib().or_ntt_batch(_ptr64(v), v.size, n)
    return v


def intt_batch(values, n: int) -> np.ndarray:
    v = u64(values).copy()
    lib().or_intt_batch(_ptr64(v), v.size, n)
    return v


def dft_naive(values) -> np.ndarray:
    v = u64(values)
    o = np.zeros_like(v)
    lib().or_dft_naive(_ptr64(v), _ptr64(o), v.size)
    return o


def rs_encode(coeffs, num_polys, poly_size, cw, depth) -> np.ndarray:
    c = u64(coeffs)
    out = np.zeros(cw * num_polys * depth, dtype=np.uint64)
    lib().or_rs_encode(_ptr64(c), num_polys, poly_size, cw, depth, _ptr64(out))
    return out


def rs_encode_ext3(coeffs, num_polys, poly_size, cw, depth) -> np.ndarray:
    c = u64(coeffs)
    out = np.zeros(3 * cw * num_polys * depth, dtype=np.uint64)
    lib().or_rs_encode_ext3(_ptr64(c), num_polys, poly_size, cw, depth, _ptr64(out))
    return out


def blake3(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    lib().or_blake3(_ptr8(buf), len(data), _ptr8(out))
    return out.tobytes()


def commit_leaves_base(matrix, rows, cols) -> np.ndarray:
    m = u64(matrix)
    out = np.zeros(32 * rows, dtype=np.uint8)
    lib().or_commit_leaves_base(_ptr64(m), rows, cols, _ptr8(out))
    return out


def commit_leaves_ext3(matrix, rows, cols) -> np.ndarray:
    m = u64(matrix)
    out = np.zeros(32 * rows, dtype=np.uint8)
    lib().or_commit_leaves_ext3(_ptr64(m), rows, cols, _ptr8(out))
    return out


# ------------------------------------------------------------ SHA-256 engine
# The second hash engine (hash/sha2_engine.hpp:62-79: SHA-256 of each message's
# bytes), restated with hashlib's SHA-256 (FIPS 180-4): row leaves over LE u64
# limbs (matrix_commit.hpp:169-195), the zero-padded tree with parent =
# SHA-256(left || right) (merkle_tree.hpp:104-147) and the PoW scan over
# SHA-256(challenge || LE64(n) || 0^24) (proof_of_work.hpp:95-193).
def sha256(data: bytes) -> bytes:
    return hashlib.sha256(bytes(data)).digest()


def sha256_leaves(matrix, rows: int, row_u64: int) -> np.ndarray:
    m = u64(matrix).reshape(rows, row_u64)
    return np.frombuffer(b"".join(sha256(m[r].tobytes()) for r in range(rows)), dtype=np.uint8).copy()


def sha256_merkle_build(leaves: np.ndarray, n: int) -> np.ndarray:
    width = 1
    while width < n:
        width <<= 1
    level = [bytes(np.asarray(leaves, dtype=np.uint8)[32 * i:32 * i + 32]) for i in range(n)]
    level += [bytes(32)] * (width - n)
    out = list(level)
    while len(level) > 1:
        level = [sha256(level[2 * i] + level[2 * i + 1]) for i in range(len(level) // 2)]
        out += level
    return np.frombuffer(b"".join(out), dtype=np.uint8).copy()


def sha256_pow_find(challenge: bytes, threshold: int) -> int:
    n = 0
    while True:
        h = sha256(bytes(challenge) + n.to_bytes(8, "little") + bytes(24))
        if int.from_bytes(h[:8], "little") <= threshold:
            return n
        n += 1


def merkle_num_nodes(n: int) -> int:
    return lib().or_merkle_num_nodes(n)


def merkle_build(leaves: np.ndarray, n: int) -> np.ndarray:
    lv = np.ascontiguousarray(leaves, dtype=np.uint8)
    if lv.size == 0:
        lv = np.zeros(32, dtype=np.uint8)
    nodes = np.zeros(32 * merkle_num_nodes(n), dtype=np.uint8)
    lib().or_merkle_build(_ptr8(lv), n, _ptr8(nodes))
    return nodes


def merkle_open(nodes: np.ndarray, n_leaves: int, idx) -> np.ndarray:
    ix = u64(idx) if len(idx) else np.zeros(1, dtype=np.uint64)
    k = len(idx)
    nh = lib().or_merkle_open(_ptr8(nodes), n_leaves, _ptr64(ix), k, None)
    out = np.zeros(32 * max(nh, 1), dtype=np.uint8)
    lib().or_merkle_open(_ptr8(nodes), n_leaves, _ptr64(ix), k, _ptr8(out))
    return out[: 32 * nh]


def merkle_verify(root: bytes, n_leaves: int, idx, leaf_hashes: np.ndarray, hints: np.ndarray) -> bool:
    r = np.frombuffer(bytes(root), dtype=np.uint8).copy()
    ix = u64(idx) if len(idx) else np.zeros(1, dtype=np.uint64)
    lh = np.ascontiguousarray(leaf_hashes, dtype=np.uint8)
    hh = np.ascontiguousarray(hints, dtype=np.uint8) if len(hints) else np.zeros(32, dtype=np.uint8)
    if lh.size == 0:
        lh = np.zeros(32, dtype=np.uint8)
    return bool(lib().or_merkle_verify(_ptr8(r), n_leaves, _ptr64(ix), _ptr8(lh), len(idx),
                                       _ptr8(hh), len(hints) // 32))


def sumcheck_compute_ext3(a, b):
    a, b = u64(a), u64(b)
    c0, c2 = np.zeros(3, dtype=np.uint64), np.zeros(3, dtype=np.uint64)
    lib().or_sumcheck_compute_ext3(_ptr64(a), _ptr64(b), a.size // 3, _ptr64(c0), _ptr64(c2))
    return c0, c2


def sumcheck_compute_base_ext3(a, b):
    a, b = u64(a), u64(b)
    c0, c2 = np.zeros(3, dtype=np.uint64), np.zeros(3, dtype=np.uint64)
    lib().or_sumcheck_compute_base_ext3(_ptr64(a), _ptr64(b), a.size, _ptr64(c0), _ptr64(c2))
    return c0, c2


def sumcheck_fold_ext3(v, r) -> np.ndarray:
    v, r = u64(v).copy(), u64(r)
    n = v.size // 3
    lib().or_sumcheck_fold_ext3(_ptr64(v), n, _ptr64(r))
    return v[: 3 * (n // 2)]


def sumcheck_fold_base_ext3(v, r) -> np.ndarray:
    v, r = u64(v), u64(r)
    out = np.zeros(3 * (v.size // 2), dtype=np.uint64)
    lib().or_sumcheck_fold_base_ext3(_ptr64(v), v.size, _ptr64(r), _ptr64(out))
    return out


def horner_base_ext3(coeffs, z) -> np.ndarray:
    c, z = u64(coeffs), u64(z)
    if c.size == 0:
        c = np.zeros(1, dtype=np.uint64)
        n = 0
    else:
        n = c.size
    o = np.zeros(3, dtype=np.uint64)
    lib().or_horner_base_ext3(_ptr64(c), n, _ptr64(z), _ptr64(o))
    return o


def horner_ext3_ext3(coeffs, z) -> np.ndarray:
    c, z = u64(coeffs), u64(z)
    o = np.zeros(3, dtype=np.uint64)
    lib().or_horner_ext3_ext3(_ptr64(c), c.size // 3, _ptr64(z), _ptr64(o))
    return o


def horner_base(coeffs, z: int) -> int:
    c = u64(coeffs)
    return lib().or_horner_base(_ptr64(c), c.size, z)


def sha3_256(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    lib().or_sha3_256(_ptr8(buf), len(data), _ptr8(out))
    return out.tobytes()


def sha3_512(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros(64, dtype=np.uint8)
    lib().or_sha3_512(_ptr8(buf), len(data), _ptr8(out))
    return out.tobytes()


def shake128(data: bytes, n: int) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros(max(n, 1), dtype=np.uint8)
    lib().or_shake128(_ptr8(buf), len(data), _ptr8(out), n)
    return out[:n].tobytes()


# ------------------------------------------------------------ prove path

def eval_eq_ext3(acc, point, scalar) -> np.ndarray:
    a = u64(acc).copy()
    pt = u64(point).reshape(-1)
    sc = u64(scalar)
    lib().or_eval_eq_ext3(_ptr64(a), _ptr64(pt), pt.size // 3, _ptr64(sc))
    return a


def geometric_accumulate_ext3(acc, points, scalars) -> np.ndarray:
    a = u64(acc).copy()
    pts, scs = u64(points).reshape(-1), u64(scalars).reshape(-1)
    lib().or_geometric_accumulate_ext3(_ptr64(a), a.size // 3, _ptr64(pts), _ptr64(scs), pts.size // 3)
    return a


def mle_evaluate(evals, point, is_base: bool) -> np.ndarray:
    e = u64(evals)
    pt = u64(point).reshape(-1)
    out = np.zeros(3, dtype=np.uint64)
    lib().or_mle_evaluate(_ptr64(e), int(is_base), _ptr64(pt), pt.size // 3, _ptr64(out))
    return out


def pow_find(challenge: bytes, threshold: int) -> int:
    ch = np.frombuffer(bytes(challenge), dtype=np.uint8).copy()
    return int(lib().or_pow_find(_ptr8(ch), threshold))


def _prm(unique=0, rate=1, iff=4, ff=4, sec=128, pow_bits=20, batch=1):
    return (ctypes.c_uint32 * 7)(unique, rate, iff, ff, sec, pow_bits, batch)


def ref_config(nv: int, rate: int, ff: int, sec: int, pow_bits: int) -> np.ndarray:
    R = ref()
    n = R.ref_whir_config(nv, rate, ff, sec, pow_bits, None, 0)
    out = np.zeros(n, dtype=np.uint64)
    R.ref_whir_config(nv, rate, ff, sec, pow_bits, _ptr64(out), n)
    return out


def ref_cli_ds(hash_id: int = 1, **prm) -> tuple[bytes, bytes]:
    pid = np.zeros(64, dtype=np.uint8)
    sid = np.zeros(32, dtype=np.uint8)
    ref().ref_cli_ds_hash(_prm(**prm), hash_id, _ptr8(pid), _ptr8(sid))
    return pid.tobytes(), sid.tobytes()


def ref_verify(nv: int, points, evaluations, narg: bytes, hints: bytes, hash_id: int = 1, **prm) -> bool:
    """The reference's own verifier (whir_verifier.hpp through the CLI's
    main.cpp:434-446 flow) on a proof of the claims f(points[e]) =
    evaluations[e * batch + v]; True when it accepts."""
    R = ref()
    pts = u64(points).reshape(-1)
    nf = pts.size // (3 * nv) if nv else 0
    ev = u64(evaluations).reshape(-1) if nf else np.zeros(3, dtype=np.uint64)
    n = np.frombuffer(bytes(narg) + b"\0", dtype=np.uint8).copy()
    h = np.frombuffer(bytes(hints) + b"\0", dtype=np.uint8).copy()
    return R.ref_whir_verify_hash(_prm(**prm), hash_id, nv, nf, _ptr64(pts) if nf else None, _ptr64(ev),
                                  _ptr8(n), len(narg), _ptr8(h), len(hints)) == 1


def ref_prove(nv: int, coeffs, points, hash_id: int = 1, **prm):
    """The reference's own prover (compiled from its sources): the CLI's PCS
    flow with caller-given coefficients (batch x 2^nv) and MLE points
    (num_constraints x nv x 3).  Returns (evaluations, narg, hints)."""
    R = ref()
    c = u64(coeffs)
    pts = u64(points).reshape(-1)
    nf = pts.size // (3 * nv) if nv else 0
    batch = prm.get("batch", 1)
    ev = np.zeros(3 * max(nf * batch, 1), dtype=np.uint64)
    cap_n, cap_h = 1 << 16, 1 << 24
    while True:
        narg = np.zeros(cap_n, dtype=np.uint8)
        hints = np.zeros(cap_h, dtype=np.uint8)
        nl, hl = _sz(), _sz()
        rc = R.ref_whir_prove_hash(_prm(**prm), hash_id, nv, _ptr64(c), nf, _ptr64(pts) if nf else None,
                                   _ptr64(ev), _ptr8(narg), cap_n, ctypes.byref(nl), _ptr8(hints), cap_h,
                                   ctypes.byref(hl))
        if rc == 0:
            return ev[:3 * nf * batch].reshape(-1, 3), narg[:nl.value].tobytes(), hints[:hl.value].tobytes()
        cap_n, cap_h = max(cap_n, nl.value), max(cap_h, hl.value)
