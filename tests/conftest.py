import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def gpu_available():
    try:
        from fibrenet_amd import lib
        import ctypes
        n = ctypes.c_int(0)
        L = ctypes.CDLL("libamdhip64.so")
        return L.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0 and lib() is not None
    except OSError:
        return False


@pytest.fixture(scope="session")
def need_gpu():
    if not gpu_available():
        pytest.fail("GPU test run without a visible HIP device")
