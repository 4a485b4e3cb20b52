import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def _gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when a gpu-marked test runs without a usable GPU."""
    if not _gpu_available():
        pytest.fail("gpu-marked test run without a visible HIP device")
    import brmerge_amd
    assert brmerge_amd.lib().brm_init(0) == 0, brmerge_amd._lib.last_error()
    return brmerge_amd
