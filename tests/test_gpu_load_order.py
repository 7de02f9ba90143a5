"""PyTorch and libbmg.so in one process, libbmg.so loaded first.

PyTorch for ROCm bundles its own HIP runtime under the soname libbmg.so
links; the runtime loaded first serves both, and torch finds no device on
the system one.  The binding loads torch first where it is installed, so the
import order of the caller does not matter.  Run in a fresh interpreter: the
test process has loaded torch already."""
import pathlib
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

CODE = r"""
import sys
import numpy as np
from briscola_amd import bmg
n = 8
verts = [np.linspace(0.0, 1.0, n + 1)] * 3
ctx = bmg.Context(verts, (bmg.DIRICHLET,) * 6, (0.0,) * 6, rel_tol=1e-8)
ctx.set_laplacian(None)
b = np.zeros((n + 2,) * 3)
b[1:-1, 1:-1, 1:-1] = 1.0
x, st = ctx.solve(b)
ctx.close()
assert st.final_residual <= 1e-8 * st.initial_residual
import torch
assert torch.cuda.is_available(), "torch sees no device"
t = torch.arange(4.0, device="cuda")
assert float(t.sum()) == 6.0
print("load order ok", flush=True)
"""


def test_torch_after_libbmg(gpu):
    print("[load order] fresh interpreter: libbmg.so, then torch", flush=True)
    r = subprocess.run([sys.executable, "-c", CODE], capture_output=True,
                       text=True, timeout=240,
                       cwd=str(pathlib.Path(__file__).parents[1]))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "load order ok" in r.stdout
