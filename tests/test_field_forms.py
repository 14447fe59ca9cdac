"""The device field forms (whirplonk_amd/csrc/gl_forms.h) built for the host
and checked against 128-bit arithmetic (tests/field_forms_check.cpp).

The GPU kernels include the same file, so this pins their add, sub, neg and
reduction on the inputs random device tests almost never produce: the
reduction's borrow cases need lo + hl*eps (mod 2^64) below a 32-bit value.
Plain Goldilocks definitions as in algebra/goldilocks.hpp:54-202.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _clang():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


def test_device_field_forms_on_host(tmp_path):
    cxx = _clang()
    if cxx is None:
        pytest.skip("no clang++ (the forms use __builtin_subc)")
    exe = str(tmp_path / "field_forms_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "whirplonk_amd", "csrc"),
                    os.path.join(ROOT, "tests", "field_forms_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
