import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running full-size check")


def gpu_available() -> bool:
    try:
        from visrtx_amd import _native
        import ctypes
        n = ctypes.c_int(0)
        return _native.hip().vrx_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_lib():
    from visrtx_amd import _native
    return _native.hip()


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
