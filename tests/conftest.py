import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def hip_runtime():
    """The HIP runtime libvlayout.so is linked against (ROCm's libamdhip64.so.7).
    Tests load this one, not a copy bundled with another package, so that
    every HIP call in the test process goes through a single runtime."""
    import ctypes as C
    return C.CDLL("libamdhip64.so.7")


def _gpu_present():
    if not os.path.exists("/dev/kfd"):
        return False
    try:
        import ctypes as C
        n = C.c_int(0)
        return hip_runtime().hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu_engine():
    if not _gpu_present():
        pytest.skip("no GPU")
    import vlsi_layout_amd as V
    return V.Engine(0)
