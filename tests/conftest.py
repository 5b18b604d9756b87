import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def _have_gpu() -> bool:
    try:
        from fluidsim_amd import _native
        import ctypes
        n = ctypes.c_int(0)
        _native.device_lib().fsa_device_count(ctypes.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_available():
    if not _have_gpu():
        pytest.fail("no gfx950 device visible or libfluidsim_amd.so not built: the HIP path must run here")
    return True
