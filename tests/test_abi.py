"""The C ABI library loads and exports every symbol include/whir_gpu.h
declares (no device work: this runs on a CPU-only machine)."""
import ctypes
import os
import re

import numpy as np
import pytest

import whirplonk_amd as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "whir_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(whir_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = W.load()
    syms = header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), s
    # the Python signature table covers exactly the header
    assert sorted(W.symbols()) == syms


def test_library_is_gfx950_and_self_contained():
    # the shared object embeds a gfx950 code object and nothing for other targets
    data = open(W.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert b"gfx942" not in data and b"gfx90a" not in data


def test_host_only_helpers(orc):
    # merkle_num_nodes and the open-path hint order are pure host functions
    for n in (0, 1, 2, 3, 17, 1024, 1 << 22):
        assert W.merkle_num_nodes(n) == orc.merkle_num_nodes(n)
    rnd = np.random.default_rng(3)
    for n in (1, 2, 3, 17, 1000):
        idx = list(rnd.integers(0, n, size=7)) + [0, n - 1, n - 1]
        leaves = rnd.integers(0, 256, size=32 * n, dtype=np.uint8)
        nodes = orc.merkle_build(leaves, n)
        hints = orc.merkle_open(nodes, n, idx)
        ni = W.merkle_hint_node_indices(n, idx)
        got = np.concatenate([nodes[32 * int(i):32 * int(i) + 32] for i in ni]) if len(ni) else np.zeros(0, np.uint8)
        assert np.array_equal(got, hints)


def test_dispatch_controls():
    lib = W.load()
    t = lib.whir_gpu_ntt_threshold()
    assert t == 65536
    lib.whir_set_gpu_ntt_threshold(123)
    assert lib.whir_gpu_ntt_threshold() == 123
    lib.whir_set_gpu_ntt_threshold(t)
    was = lib.whir_gpu_dispatch_enabled()
    lib.whir_set_gpu_dispatch_enabled(0)
    # disabled dispatch reports "not handled" (> 0) without touching a device
    v = np.arange(1 << 17, dtype=np.uint64)
    rc = lib.whir_gpu_ntt_batch(v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), v.size, 1 << 17)
    assert rc > 0
    with pytest.raises(W.WhirGpuError) as e:
        W.ntt_batch(v, 1 << 17)
    assert e.value.status > 0
    lib.whir_set_gpu_dispatch_enabled(was)


def test_no_fallback_when_library_missing(tmp_path, monkeypatch):
    import importlib

    monkeypatch.setattr(W, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(W, "_lib", None)
    with pytest.raises(ImportError):
        W.load()
    importlib.reload(W)


def test_null_handles_are_refused():
    """Entry points given a null batch, server or context report an error
    instead of touching a device."""
    lib = W.load()
    a, b = ctypes.c_double(7.0), ctypes.c_double(7.0)
    assert lib.whir_batch_form(None, ctypes.byref(a), ctypes.byref(b)) == -1
    assert lib.whir_batch_run(None) != 0
    out = np.zeros(11, dtype=np.uint64)
    assert lib.whir_server_stats(None, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) != 0
    h = ctypes.c_void_p()
    sizes = np.array([1 << 14], dtype=np.uint64)
    # a batch form outside -1..1 is rejected before any device work
    assert lib.whir_batch_create_form(None, 1, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 2, 4, 1, 0, 5,
                                      ctypes.byref(h)) != 0
    # copies through a null context, or to or from a null pointer
    buf = np.zeros(4, dtype=np.uint64)
    p = buf.ctypes.data_as(ctypes.c_void_p)
    assert lib.whir_memcpy_h2d(None, p, p, 32) != 0
    assert lib.whir_memcpy_d2h(None, p, p, 32) != 0
