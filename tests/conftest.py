import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfuss_gpu.so)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


_GPU = None


def gpu_available():
    """True when a HIP device is visible (checked through the HIP runtime,
    without importing torch)."""
    global _GPU
    if _GPU is None:
        try:
            import ctypes

            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _GPU = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle
