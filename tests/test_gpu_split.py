"""The split Poisson solver (briscola_amd/host/hipSplitPoissonSolver.H over
bmg_split_source and the FFT / multigrid solvers): the GPU source against
the CPU restatement (oracle/split_ref.py, pinned in
tests/test_split_oracle.py) bit for bit; the C++ adapter, driven the way a
flow solver drives splitPoissonSolver (splitPoissonSolver.C:53-149), equal
bit for bit to the same steps composed here from the library's own calls
(prediction, corrections, inner FFT or multigrid solve, time term, flux);
and with enough corrections the split iteration converges to the
variable-coefficient multigrid solution of the same system."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import split_ref as S
from briscola_amd import bmg, workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_split")


def metrics(prob):
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    w = [o.widths(0, d) for d in range(3)]
    dl = []
    for d in range(3):
        cc = o.centres(0, d)
        a = np.zeros_like(cc)
        a[1:] = 1.0 / np.abs(cc[1:] - cc[:-1])
        dl.append(a)
    return w, dl


def face_max(lam, n):
    m = 0.0
    for d in range(3):
        hi = [q + 1 for q in n]
        hi[d] += 1
        m = max(m, lam[d][tuple(slice(1, h) for h in hi)].max())
    return m


# nonzero Dirichlet values (y walls) and Neumann gradients (z walls) of the
# "smooth" problem's faces
BC_VALUES = (0.0, 0.0, 0.5, -1.0, 0.3, -0.2)


def problem(case, n=16, bc_value=(0.0,) * 6):
    rng = np.random.default_rng(21)
    if case == "spheres":
        p = W.spheres(n, seed=2, nspheres=3, rho_out=10.0)
        lam = p.lam
    else:
        verts = [np.linspace(0, 1, n + 1)] * 3
        bc = (W.PERIODIC, W.PERIODIC, W.DIRICHLET, W.DIRICHLET,
              W.NEUMANN, W.NEUMANN)
        b = np.zeros((n + 2,) * 3)
        b[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (n,) * 3)
        p = W.Problem("smooth", verts, bc, b, bc_value=bc_value)
        c = [np.linspace(0, 1, n + 2)] * 3
        base = 0.75 + 0.25 * np.sin(2 * np.pi * c[0][:, None, None]) * \
            np.cos(np.pi * c[1][None, :, None]) * np.ones((1, 1, n + 2))
        lam = [base.copy() for _ in range(3)]
    return p, [np.ascontiguousarray(a) for a in lam]


@pytest.mark.parametrize("case", ["spheres", "smooth"])
def test_split_source_matches_restatement(gpu, case):
    prob, lam = problem(case)
    print(f"[split source] case={case}", flush=True)
    w, dl = metrics(prob)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(prob.padded)
    l0 = face_max(lam, prob.n)
    ctx = gpu.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-8)
    for b in (prob.b, None):
        bh, ph = ctx.split_source(x, lam, l0, b)
        rbh, rph = S.source(x, lam, l0, b, w, dl)
        assert np.array_equal(bh, rbh)
        for d in range(3):
            assert np.array_equal(ph[d], rph[d])
    ctx.close()


def run_adapter(tmp_path, prob, lam, inner, ncorr, prediction, ddt, x0, xold,
                xoldold, rel_tol="1e-10", precond=False):
    dt, dt0, frac = ddt if ddt is not None else (1e-3, 1e-3, 1.0)
    with open(tmp_path / "shape.txt", "w") as f:
        f.write(" ".join(str(q) for q in prob.n) + " ")
        f.write(" ".join(str(q) for q in prob.bc) +
                f" {inner} {ncorr} {'true' if prediction else 'false'} "
                f"{int(ddt is not None)} {dt!r} {dt0!r} {frac!r} {rel_tol} "
                f"{int(precond)} " +
                " ".join(repr(float(v)) for v in prob.bc_value) + "\n")
    w, _ = metrics(prob)
    V = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
    np.concatenate(prob.vertices).tofile(tmp_path / "vertices.bin")
    V.tofile(tmp_path / "volumes.bin")
    np.concatenate([a.ravel() for a in lam]).tofile(tmp_path / "lambda.bin")
    prob.b.tofile(tmp_path / "src.bin")
    x0.tofile(tmp_path / "x0.bin")
    xold.tofile(tmp_path / "xold.bin")
    xoldold.tofile(tmp_path / "xoldold.bin")
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    x = np.fromfile(tmp_path / "x.bin").reshape(prob.padded)
    bh = np.fromfile(tmp_path / "bhat.bin").reshape(prob.padded)
    fl = np.fromfile(tmp_path / "flux.bin").reshape((3,) + prob.padded)
    return x, bh, fl, V


@pytest.mark.parametrize("inner", ["FFT", "MG"])
@pytest.mark.parametrize("mode", ["plain", "prediction_ddt", "precond_ddt"])
def test_split_adapter_equals_composed_steps(gpu, tmp_path, inner, mode):
    prob, lam = problem("spheres")
    print(f"[split adapter] inner={inner} mode={mode}", flush=True)
    rng = np.random.default_rng(5)
    x0 = np.zeros(prob.padded)
    xold = rng.standard_normal(prob.padded) * 1e-3
    xoldold = rng.standard_normal(prob.padded) * 1e-3
    prediction = mode != "plain"
    precond = mode == "precond_ddt"
    ddt = (2e-3, 4e-3, 0.5) if mode != "plain" else None
    ncorr = 3
    x, bh, fl, V = run_adapter(tmp_path, prob, lam, inner, ncorr, prediction,
                               ddt, x0, xold, xoldold, precond=precond)
    # the same steps composed from the library's calls
    l0 = face_max(lam, prob.n)
    xr = x0.copy()
    if prediction:
        r = (ddt[0] * ddt[2]) / ddt[1]
        xr = (1.0 + r) * xold - r * xoldold
    if precond:
        pre = gpu.Context(prob.vertices, prob.bc, prob.bc_value,
                          rel_tol=1e-10, tolerance=0.0, n_sweeps_pre=2,
                          n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_level=1)
        pre.set_laplacian(lam, ddt=(ddt[0], ddt[2]))
        bs = 0.0 - V * prob.b
        bs = bs - ((V * xold) / ddt[0]) / ddt[2]
        xr, _ = pre.solve(bs, xr)
        pre.close()
    geo = gpu.Context(prob.vertices, prob.bc, prob.bc_value)
    if inner == "FFT":
        solver = gpu.FFT(prob.vertices, prob.bc, prob.bc_value)
    else:
        solver = gpu.Context(prob.vertices, prob.bc, prob.bc_value,
                             rel_tol=1e-10, tolerance=0.0, n_sweeps_pre=2,
                             n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_level=1)
        solver.set_laplacian(None, ddt=None if ddt is None else (ddt[0], ddt[2]))
    for corr in range(ncorr):
        bhr, ph = geo.split_source(xr, lam, l0, prob.b)
        if inner == "FFT":
            solver.solve(bhr, xr, ddt=None if ddt is None else (ddt[0], ddt[2]),
                         x_old=xold if ddt is not None else None)
        else:
            bs = 0.0 - V * bhr
            if ddt is not None:
                bs = bs - ((V * xold) / ddt[0]) / ddt[2]
            # the inner defaultPoissonSolver assembles its system anew
            solver.renew_source_terms()
            xr, _ = solver.solve(bs, xr)
    assert np.array_equal(bh, bhr)
    assert np.array_equal(x, xr)
    f = geo.face_flux(xr)
    for d in range(3):
        assert np.array_equal(fl[d], l0 * (f[d] + ph[d]))
    geo.close()
    solver.close()


def test_split_iteration_converges_to_variable_coefficient_solution(gpu, tmp_path):
    """Corrections with the FFT inner solver converge to the solution of
    div(lambda grad x) = -b, the multigrid's system with lambda (uniform
    mesh: the FFT's finite differences are the finite volumes over V)."""
    prob, lam = problem("smooth", n=16)
    z = np.zeros(prob.padded)
    x, _, _, V = run_adapter(tmp_path, prob, lam, "FFT", 80, False, None, z,
                             z, z)
    ctx = gpu.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-12,
                      tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
                      cycle=bmg.VCYCLE, coarse_level=1, max_iter=200)
    ctx.set_laplacian(lam)
    xm, st = ctx.solve(0.0 - V * prob.b)
    ctx.close()
    xi, mi = x[1:-1, 1:-1, 1:-1], xm[1:-1, 1:-1, 1:-1]
    err = np.abs(xi - mi).max() / np.abs(mi).max()
    assert err < 1e-8, err


def test_split_source_with_boundary_data(gpu):
    """Nonzero Dirichlet values and Neumann gradients: the source reads the
    ghosts of x, here set by the oracle's correctBoundaryConditions from
    that data, bitwise against the restatement."""
    prob, lam = problem("smooth", bc_value=BC_VALUES)
    w, dl = metrics(prob)
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    x = np.random.default_rng(4).standard_normal(prob.padded)
    o.set_field(oracle.X, 0, x)
    o.correct_bcs(0)
    x = o.field(oracle.X, 0)
    assert not np.array_equal(x[1:-1, 0, 1:-1], -x[1:-1, 1, 1:-1])
    l0 = face_max(lam, prob.n)
    ctx = gpu.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-8)
    bh, ph = ctx.split_source(x, lam, l0, prob.b)
    rbh, rph = S.source(x, lam, l0, prob.b, w, dl)
    assert np.array_equal(bh, rbh)
    for d in range(3):
        assert np.array_equal(ph[d], rph[d])
    ctx.close()


@pytest.mark.parametrize("inner", ["FFT", "MG"])
def test_split_adapter_with_boundary_data(gpu, tmp_path, inner):
    """The adapter with nonzero boundary data equals the composed steps
    bitwise (every inner multigrid solve a newly assembled system)."""
    prob, lam = problem("smooth", bc_value=BC_VALUES)
    print(f"[split adapter bc] inner={inner}", flush=True)
    z = np.zeros(prob.padded)
    x, bh, fl, V = run_adapter(tmp_path, prob, lam, inner, 3, False, None, z,
                               z, z)
    l0 = face_max(lam, prob.n)
    geo = gpu.Context(prob.vertices, prob.bc, prob.bc_value)
    if inner == "FFT":
        solver = gpu.FFT(prob.vertices, prob.bc, prob.bc_value)
    else:
        solver = gpu.Context(prob.vertices, prob.bc, prob.bc_value,
                             rel_tol=1e-10, tolerance=0.0, n_sweeps_pre=2,
                             n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_level=1)
        solver.set_laplacian(None)
    xr = z.copy()
    for corr in range(3):
        bhr, ph = geo.split_source(xr, lam, l0, prob.b)
        if inner == "FFT":
            solver.solve(bhr, xr)
        else:
            solver.renew_source_terms()
            xr, _ = solver.solve(0.0 - V * bhr, xr)
    assert np.array_equal(bh, bhr)
    assert np.array_equal(x, xr)
    geo.close()
    solver.close()


@pytest.mark.parametrize("inner", ["FFT", "MG"])
def test_split_with_boundary_data_converges(gpu, tmp_path, inner):
    """With nonzero Dirichlet values and Neumann gradients the split
    iteration (FFT or multigrid inner solves) converges to the multigrid
    solution of div(lambda grad x) = -b with the same boundary data, which
    is itself bitwise the oracle's."""
    prob, lam = problem("smooth", n=16, bc_value=BC_VALUES)
    z = np.zeros(prob.padded)
    x, _, _, V = run_adapter(tmp_path, prob, lam, inner, 80, False, None, z,
                             z, z, rel_tol="1e-12")
    ctx = gpu.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-12,
                      tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
                      cycle=bmg.VCYCLE, coarse_level=1, max_iter=200)
    ctx.set_laplacian(lam)
    xm, st = ctx.solve(0.0 - V * prob.b)
    ctx.close()
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    o.set_opts(rel_tol=1e-12, tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
               cycle=oracle.VCYCLE, coarse_level=1, max_iter=200)
    o.assemble(lam)
    o.set_field(oracle.B, 0, 0.0 - V * prob.b)
    o.solve()
    assert np.array_equal(xm, o.field(oracle.X, 0))
    xi, mi = x[1:-1, 1:-1, 1:-1], xm[1:-1, 1:-1, 1:-1]
    err = np.abs(xi - mi).max() / np.abs(mi).max()
    print(f"[split bc] inner={inner}: 80 corrections, error vs MG {err:.2e}",
          flush=True)
    assert err < 1e-8, err
