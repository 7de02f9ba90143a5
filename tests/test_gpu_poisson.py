"""The PoissonSolver adapter (briscola_amd/host/hipPoissonSolver.H) driven
by a C++ program the way a Briscola flow solver drives defaultPoissonSolver
(defaultPoissonSolver.C:39-124): source term b_sys = 0 - V*s, GPU assembly
of laplacian(lambda, x), Runge-Kutta stage warm starts and the face flux.
Every solve must equal a direct bmg_solve of the same system from the
initial guess the stage logic prescribes, bitwise, and the flux must equal
the reference formula (lambda*((x_lower - x)*delta))*area evaluated here
with numpy in the same order, bitwise."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from briscola_amd import bmg, workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_poisson")
STAGES = 3


def metrics(prob):
    """Widths and lower-face 1/delta per padded index, from the oracle's
    level-0 geometry (the solver's own construction)."""
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    w = [o.widths(0, d) for d in range(3)]
    dl = []
    for d in range(3):
        cc = o.centres(0, d)
        a = np.zeros_like(cc)
        a[1:] = 1.0 / np.abs(cc[1:] - cc[:-1])
        dl.append(a)
    return w, dl


def flux_ref(x, lam, w, dl, n):
    """defaultPoissonSolver.C:108-121 face by face, numpy, same order."""
    out = []
    for d in range(3):
        f = np.zeros_like(x)
        hi = [q + 1 for q in n]
        hi[d] += 1
        sl = tuple(slice(1, h) for h in hi)
        lo = tuple(slice(1 - (e == d), h - (e == d)) for e, h in enumerate(hi))
        g = (x[lo] - x[sl])
        idx = [np.arange(1, h) for h in hi]
        dd = dl[d][idx[d]].reshape([-1 if e == d else 1 for e in range(3)])
        g = g * dd
        a, b = [e for e in range(3) if e != d]
        fa = (w[a][idx[a]].reshape([-1 if e == a else 1 for e in range(3)])
              * w[b][idx[b]].reshape([-1 if e == b else 1 for e in range(3)]))
        v = lam[d][sl] * g if lam is not None else g
        f[sl] = v * fa
        out.append(f)
    return out


@pytest.mark.parametrize("case", ["spheres", "channel"])
def test_poisson_solver_adapter(gpu, tmp_path, case):
    if case == "spheres":
        prob = W.spheres(16, seed=3, nspheres=4)
        lam = prob.lam
    else:
        prob = W.channel(16, seed=5)
        lam = None
    print(f"[poisson] case={case}", flush=True)
    n = prob.n
    w, dl = metrics(prob)
    V = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
    rng = np.random.default_rng(7)
    src = [np.zeros(prob.padded) for _ in range(STAGES + 1)]
    for s in src:
        s[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, n)
        if W.DIRICHLET not in prob.bc:
            inner = s[1:-1, 1:-1, 1:-1]
            inner -= (V[1:-1, 1:-1, 1:-1] * inner).sum() / V[1:-1, 1:-1, 1:-1].sum()
    with open(tmp_path / "shape.txt", "w") as f:
        f.write(" ".join(str(q) for q in n) + " ")
        f.write(" ".join(str(q) for q in prob.bc) + f" 1e-9 1 {STAGES} "
                f"{int(lam is not None)}\n")
    np.concatenate(prob.vertices).tofile(tmp_path / "vertices.bin")
    V.tofile(tmp_path / "volumes.bin")
    if lam is not None:
        np.concatenate([a.ravel() for a in lam]).tofile(tmp_path / "lambda.bin")
    np.concatenate([s.ravel() for s in src]).tofile(tmp_path / "src.bin")
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("MG: Solving for colocated p, initial residual")

    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-9,
                      tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
                      cycle=bmg.VCYCLE, coarse_level=1)
    ctx.set_laplacian(lam)
    stats = [l.split() for l in open(tmp_path / "stats.txt")]
    stages = [np.zeros(prob.padded) for _ in range(STAGES)]
    for k in range(STAGES + 1):
        stage = k if k < STAGES else 0
        b = 0.0 - V * src[k]
        x0 = stages[stage].copy()
        x, st = ctx.solve(b, x0)
        stages[stage] = x.copy()
        xg = np.fromfile(tmp_path / f"x{k}.bin").reshape(prob.padded)
        assert np.array_equal(xg, x), f"solve {k}"
        assert int(stats[k][0]) == st.iters
        fl = np.fromfile(tmp_path / f"flux{k}.bin").reshape((3,) + prob.padded)
        ref = flux_ref(xg, lam, w, dl, n)
        for d in range(3):
            assert np.array_equal(fl[d], ref[d]), f"solve {k} flux {d}"
    ctx.close()
