import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libtangle_hip.so")


@pytest.fixture(scope="session")
def hip_device():
    """Fails loudly when a GPU test runs without the HIP library or a device."""
    from tangle_amd import _native

    lib = _native.library()  # raises RuntimeError when the library is not built
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    count = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(count)) != 0 or count.value < 1:
        pytest.fail("gpu test selected but no HIP device is visible")
    return lib
