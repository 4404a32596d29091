import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "ref: needs the reference CPU build under oracle/_ref")


@pytest.fixture(scope="session")
def gpu():
    """Initialise device 0 once; fail loudly when the HIP library or the GPU is missing."""
    from niftyreg_amd import _lib
    import ctypes
    lib = _lib.load()
    n = ctypes.c_int(0)
    _lib.check(lib.nrh_device_count(ctypes.byref(n)))
    if n.value < 1:
        pytest.fail("no HIP device visible")
    _lib.check(lib.nrh_init(0))
    return lib
