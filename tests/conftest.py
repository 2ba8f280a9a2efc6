import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libaxicor_day.so")
    config.addinivalue_line("markers", "slow: long CPU-side test")


def gpu_available() -> bool:
    """A ROCm device node is present (no GPU call is made here)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from axicor_amd._ffi import day_lib
    lib = day_lib()
    lib.gpu_set_device(0)
    return lib
