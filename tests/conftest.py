import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gpu():
    from briscola_amd import bmg
    bmg.load()
    import ctypes
    n = ctypes.c_int(0)
    code = bmg.load().bmg_device_count(ctypes.byref(n))
    if code != 0 or n.value < 1:
        pytest.fail("no HIP device visible")
    return bmg
