import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "generalized-toric-gpu-search_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def gpu_ctx():
    from gtoric import _native
    return _native.context(0)
