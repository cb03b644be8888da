import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libns3d_gpu.so")


def _gpu_available():
    if not os.path.exists("/dev/kfd"):
        return False
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip cleanly when no device; fail loudly when the library is missing."""
    if not _gpu_available():
        pytest.skip("no HIP device")
    import ns3d_cpp_amd
    ns3d_cpp_amd.lib()  # raises if libns3d_gpu.so is missing
    return True
