import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available():
    try:
        import ctypes
        from polymage_amd import runtime
        n = ctypes.c_int(0)
        return runtime.rt().pmx_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if os.environ.get('PMX_PREBUILD') or os.environ.get('PMX_EMULATE', '') not in ('', '0'):
        return True
    if not gpu_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return True
