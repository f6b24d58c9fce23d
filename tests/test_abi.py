"""The C-ABI library loads and exports every function include/tnad.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "tnad.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tnad_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ["tnad_create", "tnad_ctmrg", "tnad_ctmrg_back", "tnad_energy", "tnad_trg", "tnad_ctmrgstep",
                 "tnad_symeig", "tnad_svd_back_sym", "tnad_expectationvalue"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import tnad._lib as L
    if not L.LIB_PATH.exists():
        pytest.fail("libtnad.so is not built (run tensornetworkad.jl_amd/build_lib.py)")
    lib = ctypes.CDLL(str(L.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes layer declares exactly the header's functions
    assert sorted(L.SIGNATURES) == declared_functions()


def test_ctypes_layer_loads_and_reports_version():
    import tnad._lib as L
    lib = L.load()
    assert lib.tnad_version() == 1
    o = L.Opts()
    lib.tnad_opts_default(ctypes.byref(o))
    assert o.precision == 0 and o.eig_max_iters == 200 and o.eig_tol == pytest.approx(1e-13)


def test_no_cpu_fallback_in_product_path():
    # the host layer never imports the oracle
    pkg = ROOT / "tensornetworkad.jl_amd"
    for p in pkg.rglob("*.py"):
        src = p.read_text()
        assert "oracle" not in src.replace("oracle.", "").lower() or p.name == "build_lib.py", p


def test_trg_svd_back_checks_cotangent_shapes_before_the_library():
    # host-side validation: a malformed cotangent never reaches the C ABI (no context is needed)
    import numpy as np
    import tnad
    t = np.zeros((2, 3, 3, 1))
    with pytest.raises(ValueError):
        tnad.trg_svd_back(t, 10, 0.0, np.zeros((6, 3)), np.zeros((3, 3, 1)), ctx=object())
    with pytest.raises(ValueError):
        tnad.trg_svd_back(t, 10, 0.0, np.zeros((2, 3, 3)), np.zeros((2, 3, 1)), ctx=object())
