import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("native-firesim_amd", "oracle", "tests"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfiresim.so")


def gpu_available() -> bool:
    """True when the HIP runtime reports at least one device."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return False
    count = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(count)) == 0 and count.value > 0


@pytest.fixture(scope="session")
def hip_lib():
    import firesim
    if not gpu_available():
        pytest.fail("GPU test collected on a host without a HIP device")
    return firesim.load_library()
