import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libdvide_hip.so")


def _gpu_available():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a GPU test fails loudly where no GPU is visible instead of skipping."""
    if not _gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
    from dvide_amd import system
    system.load_library()
    return True
