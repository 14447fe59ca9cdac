"""The library's host concurrency helpers (whirplonk_amd/csrc/hostsync.h: the
copy thread pool of the staged host copies and the wait hub of the serving
loop) built for the host alone and run under stress (tests/hostsync_check.cpp):
with ThreadSanitizer (ROCm's clang, whose runtime knows the clock-based
condition-variable waits that std::condition_variable::wait_for uses), and
without it.  No GPU is involved; a hang (a lost wake-up, a job that never
completes) fails the test through the time limit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.parametrize("tsan", [True, False])
def test_host_sync_helpers(tmp_path, tsan):
    cxx = CLANG if tsan else shutil.which("g++")
    if cxx is None or not os.path.exists(cxx):
        pytest.skip("no compiler for this build")
    exe = str(tmp_path / ("hostsync_tsan" if tsan else "hostsync"))
    flags = ["-O1", "-g", "-fsanitize=thread"] if tsan else ["-O2"]
    subprocess.run([cxx, "-std=c++17", "-pthread", *flags, "-I", os.path.join(ROOT, "whirplonk_amd", "csrc"),
                    os.path.join(ROOT, "tests", "hostsync_check.cpp"), "-o", exe], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert "ThreadSanitizer" not in out.stderr, out.stderr
