"""The C++ adapter (briscola_amd/host/hipMG.H) driven by a C++ program the
way Briscola drives MG::solve(linearSystem&, bool), and the diagonal-system
path, both compared bit for bit with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from briscola_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_hipmg")


@pytest.mark.parametrize("case", ["neumann", "channel"])
def test_cpp_adapter_solve(gpu, tmp_path, case):
    prob = W.uniform_neumann(32, seed=4) if case == "neumann" else W.channel(16, seed=4)
    print(f"[adapter] case={case}", flush=True)
    o = oracle.Oracle(prob.vertices, prob.bc)
    o.set_opts(rel_tol=1e-9, tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
               cycle=oracle.VCYCLE, coarse_level=1, max_iter=999)
    o.assemble()
    with open(tmp_path / "shape.txt", "w") as f:
        f.write(" ".join(str(q) for q in prob.n) + f" {o.nlevels} ")
        f.write(" ".join(str(q) for q in prob.bc) + " 1e-9 1\n")
        for s in o.shapes:
            f.write(" ".join(str(q) for q in s) + "\n")
    np.concatenate(prob.vertices).tofile(tmp_path / "vertices.bin")
    for l in range(o.nlevels):
        o.stencil(l).tofile(tmp_path / f"A{l}.bin")
    prob.b.tofile(tmp_path / "b.bin")
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    line = out.stdout.strip().splitlines()[0]
    assert line.startswith("MG: Solving for colocated p, initial residual = ")
    o.set_field(oracle.B, 0, prob.b)
    it, r0, r1 = o.solve()
    st = open(tmp_path / "stats.txt").read().split()
    assert int(st[0]) == it
    shp = o.padded(0)
    x = np.fromfile(tmp_path / "x.bin").reshape(shp)
    assert np.array_equal(x, o.field(oracle.X, 0))
    A0 = np.fromfile(tmp_path / "A0_out.bin").reshape(shp + (7,))
    # the caller's A comes back eliminated, as MG::solve leaves it
    assert np.array_equal(A0[1:-1, 1:-1, 1:-1], o.stencil(0)[1:-1, 1:-1, 1:-1])


def test_diagonal_system(gpu):
    """MG.C:404-428: x = b / A_C, then the level-0 boundary update."""
    n = 12
    v = [np.linspace(0, 1, n + 1)] * 3
    bc = (W.DIRICHLET, W.NEUMANN, W.PERIODIC, W.PERIODIC, W.NEUMANN, W.DIRICHLET)
    rng = np.random.default_rng(2)
    A = np.zeros((n + 2,) * 3 + (7,))
    A[..., 0] = rng.uniform(1.0, 2.0, (n + 2,) * 3)
    b = rng.uniform(-1, 1, (n + 2,) * 3)
    o = oracle.Oracle(v, bc, (0.5, 0, 0, 0, 0, 0.25))
    o.set_stencil(0, A)
    o.set_field(oracle.B, 0, b)
    o.set_field(oracle.X, 0, np.zeros((n + 2,) * 3))
    o.solve_diagonal()
    ctx = gpu.Context(v, bc, (0.5, 0, 0, 0, 0, 0.25))
    bb = b.copy()
    x = np.zeros((n + 2,) * 3)
    ctx.solve_diagonal(A, bb, x)
    assert np.array_equal(x, o.field(oracle.X, 0))
    assert np.array_equal(bb, o.field(oracle.B, 0))
