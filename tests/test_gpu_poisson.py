"""The PoissonSolver adapter (briscola_amd/host/hipPoissonSolver.H) driven
by a C++ program the way a Briscola flow solver drives defaultPoissonSolver
(defaultPoissonSolver.C:39-124): source term b_sys = 0 - V*s, GPU assembly
of laplacian(lambda, x), optionally minus the implicit Euler term
ddt(x)/dtFrac, Runge-Kutta stage warm starts and the face flux.  Every solve
must equal a direct bmg_solve of the same system from the initial guess the
stage logic prescribes, bitwise, and the flux must equal the reference
formula (lambda*((x_lower - x)*delta))*area evaluated here with numpy in
the same order, bitwise.  With ddt and no source one solve is a backward
Euler step, which damps a Dirichlet sine mode by the closed-form factor."""
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


def write_inputs(tmp_path, prob, lam, V, src, ddt, rel_tol="1e-9",
                 coarse_level=1, xold=None):
    dt, frac = ddt if ddt is not None else (0.0, 1.0)
    with open(tmp_path / "shape.txt", "w") as f:
        f.write(" ".join(str(q) for q in prob.n) + " ")
        f.write(" ".join(str(q) for q in prob.bc) +
                f" {rel_tol} {coarse_level} {len(src) - 1} "
                f"{int(lam is not None)} {int(ddt is not None)} "
                f"{dt!r} {frac!r} " +
                " ".join(repr(float(v)) for v in prob.bc_value) + "\n")
    np.concatenate(prob.vertices).tofile(tmp_path / "vertices.bin")
    V.tofile(tmp_path / "volumes.bin")
    if lam is not None:
        np.concatenate([a.ravel() for a in lam]).tofile(tmp_path / "lambda.bin")
    np.concatenate([s.ravel() for s in src]).tofile(tmp_path / "src.bin")
    if xold is not None:
        xold.tofile(tmp_path / "xold.bin")


def run_driver(tmp_path):
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("MG: Solving for colocated p, initial residual")


DDT = (2e-3, 0.5)


def _with_values(prob, bc, values):
    prob.bc = tuple(bc)
    prob.bc_value = tuple(float(v) for v in values)
    return prob


@pytest.mark.parametrize("case", ["spheres", "channel", "spheres_ddt",
                                  "channel_ddt", "channel_bc", "mixed_bc",
                                  "mixed_bc_ddt"])
def test_poisson_solver_adapter(gpu, tmp_path, case):
    """Inhomogeneous cases (_bc): Neumann gradients (antisymmetric on the
    walls of the singular channel) and Dirichlet values; the adapter builds
    the system anew for every solve, so every solve, also with the matrix
    kept, folds them into b -- each solve is also checked bitwise against
    a freshly assembled oracle system."""
    if case.startswith("spheres"):
        prob = W.spheres(16, seed=3, nspheres=4)
        lam = prob.lam
    elif case.startswith("mixed"):
        prob = _with_values(W.channel(16, seed=6),
                            (W.PERIODIC, W.PERIODIC, W.DIRICHLET, W.NEUMANN,
                             W.PERIODIC, W.PERIODIC), (0, 0, 1.5, -0.75, 0, 0))
        lam = None
    else:
        prob = W.channel(16, seed=5)
        if case.endswith("_bc"):
            prob = _with_values(prob, prob.bc, (0, 0, -0.8, 0.8, 0, 0))
        lam = None
    ddt = DDT if case.endswith("_ddt") else None
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
    write_inputs(tmp_path, prob, lam, V, src, ddt)
    run_driver(tmp_path)

    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-9,
                      tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
                      cycle=bmg.VCYCLE, coarse_level=1)
    ctx.set_laplacian(lam, ddt=ddt)
    stats = [l.split() for l in open(tmp_path / "stats.txt")]
    stages = [np.zeros(prob.padded) for _ in range(STAGES)]
    xold = np.zeros(prob.padded)
    inhom = any(v != 0.0 for v in prob.bc_value)
    for k in range(STAGES + 1):
        stage = k if k < STAGES else 0
        b = 0.0 - V * src[k]
        if ddt is not None:
            b = b - ((V * xold) / ddt[0]) / ddt[1]
        x0 = stages[stage].copy()
        if inhom:
            o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
            o.set_opts(rel_tol=1e-9, tolerance=0.0, n_sweeps_pre=2,
                       n_sweeps_post=2, cycle=oracle.VCYCLE, coarse_level=1)
            o.assemble(lam, ddt=ddt)
            o.set_field(oracle.B, 0, b)
            o.set_field(oracle.X, 0, x0)
            o.solve()
        ctx.renew_source_terms()
        x, st = ctx.solve(b, x0)
        stages[stage] = x.copy()
        xold = x.copy()
        xg = np.fromfile(tmp_path / f"x{k}.bin").reshape(prob.padded)
        assert np.array_equal(xg, x), f"solve {k}"
        if inhom:
            assert np.array_equal(x, o.field(oracle.X, 0)), f"oracle {k}"
        assert int(stats[k][0]) == st.iters
        fl = np.fromfile(tmp_path / f"flux{k}.bin").reshape((3,) + prob.padded)
        ref = flux_ref(xg, lam, w, dl, n)
        for d in range(3):
            assert np.array_equal(fl[d], ref[d]), f"solve {k} flux {d}"
    ctx.close()


def test_poisson_solver_ddt_is_backward_euler(gpu, tmp_path):
    """No source, x.oldTime() a Dirichlet sine mode: each solve is one
    backward Euler step of V dx/dt = laplacian(x) with step dt*dtFrac, so
    x_k = g^(k+1) x_old with g = 1/(1 - dt dtFrac lambda/V), lambda the
    mode's eigenvalue of the 7-point operator."""
    n = 16
    verts = [np.linspace(0, 1, n + 1)] * 3
    prob = W.Problem("dir", verts, (W.DIRICHLET,) * 6, np.zeros((n + 2,) * 3))
    h = 1.0 / n
    V = np.full(prob.padded, h ** 3)
    modes = (2, 3, 1)
    ax = [np.sin(np.pi * m * (np.arange(n + 2) - 0.5) / n) for m in modes]
    xold = ax[0][:, None, None] * ax[1][None, :, None] * ax[2][None, None, :]
    dt, frac = 1e-2, 0.5
    src = [np.zeros(prob.padded) for _ in range(2)]
    write_inputs(tmp_path, prob, None, V, src, (dt, frac), rel_tol="1e-12",
                 xold=xold)
    run_driver(tmp_path)
    lam = sum(h * (2 * np.cos(np.pi * m / n) - 2) for m in modes)
    g = 1.0 / (1.0 - dt * frac * lam / h ** 3)
    assert g < 0.7
    for k in range(2):
        x = np.fromfile(tmp_path / f"x{k}.bin").reshape(prob.padded)
        ref = g ** (k + 1) * xold
        err = np.abs(x[1:-1, 1:-1, 1:-1] - ref[1:-1, 1:-1, 1:-1]).max()
        assert err <= 1e-9 * np.abs(xold).max(), (k, err)
