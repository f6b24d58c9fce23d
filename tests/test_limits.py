"""The host-side size limits (csrc/tnad_limits.h) that every entry point checks before it launches a kernel:
compiled with g++ alone (no HIP) and exercised at the boundaries of each limit."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_size_limits(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "limits_check"
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-Werror", "-I",
                        str(ROOT / "tensornetworkad.jl_amd" / "csrc"),
                        str(ROOT / "tests" / "cpp" / "limits_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 bad" in r.stdout
