import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "tensornetworkad.jl_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libtnad.so")


@pytest.fixture(scope="session")
def ctx():
    import tnad
    c = tnad.Context(0)
    yield c
    c.close()
