import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP kernels through the C ABI")


def _device_count():
    from fastwigner_amd import _lib
    lib = _lib.load()   # raises if the library is missing: no silent fallback
    import ctypes
    n = ctypes.c_int(0)
    lib.fw_device_count(ctypes.byref(n))
    return n.value


@pytest.fixture(scope="session")
def gpu():
    """Skip when no GPU is visible; fail loudly when the native library is missing."""
    if _device_count() == 0:
        pytest.skip("no GPU visible")
    return 0
