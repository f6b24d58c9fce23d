"""The einsum-string front end (csrc/tc_build.cpp) produces descriptors that
evaluate every contraction of the CTMRG step, its adjoint and the energy chain
exactly as a direct index-by-index einsum does (host-only C++ harness)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_contraction_descriptors(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "tc_build_check"
    r = subprocess.run([gxx, "-O1", "-std=c++17", "-I", str(ROOT / "tensornetworkad.jl_amd" / "csrc"),
                        str(ROOT / "tests" / "cpp" / "tc_build_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 bad" in r.stdout
