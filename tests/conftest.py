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
    # torch (device buffers of the device-pointer tests) ships its own HIP
    # runtime under the same soname as ROCm's: load it before the solver
    # library so that one runtime, torch's, serves both
    try:
        import torch
    except ImportError:
        torch = None
    from briscola_amd import bmg
    bmg.load()
    import ctypes
    n = ctypes.c_int(0)
    code = bmg.load().bmg_device_count(ctypes.byref(n))
    if code != 0 or n.value < 1:
        pytest.fail("no HIP device visible")
    if torch is not None and torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")
    return bmg
