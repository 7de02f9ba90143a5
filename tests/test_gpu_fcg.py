"""Flexible CG with an MG-cycle preconditioner (FP64 or mixed precision).

Not a reference algorithm, so it is pinned by what the mathematics fixes:
the returned x satisfies the reference stopping test on the true residual,
matches an FP64 direct solve of the same assembled system, and needs fewer
iterations than plain MG cycles on the anisotropic channel."""
import numpy as np
import pytest

import oracle
from oracle import direct
from briscola_amd import bmg, workloads as W

pytestmark = pytest.mark.gpu

BASE = dict(rel_tol=1e-8, tolerance=0.0, max_iter=999, n_sweeps_pre=2,
            n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_mode=bmg.COARSE_DIRECT,
            coarse_level=1)


def solve(prob, **kw):
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, **dict(BASE, **kw))
    ctx.set_laplacian(prob.lam)
    b = prob.b.copy()
    x, st = ctx.solve(b)
    return x, st


def exact(prob):
    if prob.name == "channel":
        return direct.solve_fft_thomas(prob.b, prob.vertices)
    o = oracle.Oracle(prob.vertices, prob.bc)
    o.assemble(prob.lam)
    o.pin_and_eliminate()
    return direct.solve_sparse(o.stencil(0), prob.b, prob.bc)


@pytest.mark.parametrize("precision", [bmg.FP64, bmg.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", ["channel32", "spheres32", "neumann32"])
def test_fcg_solves_to_tolerance(gpu, case, precision):
    prob = {"channel32": lambda: W.channel(32, seed=6),
            "spheres32": lambda: W.spheres(32, seed=6, nspheres=8),
            "neumann32": lambda: W.uniform_neumann(32, seed=6)}[case]()
    print(f"[fcg] case={case} precision={precision}", flush=True)
    x, st = solve(prob, precision=precision, krylov=bmg.KRYLOV_FCG)
    assert st.final_residual <= 1e-8 * st.initial_residual
    xe = exact(prob)
    err = direct.rel_error(x[1:-1, 1:-1, 1:-1], xe, True)
    assert err < 1e-6, err
    xm, stm = solve(prob, precision=bmg.FP64)
    assert st.iters <= stm.iters
    if case == "channel32":
        assert st.iters * 3 < stm.iters


def test_fcg_dirichlet_nonsingular(gpu):
    n = 24
    v = [np.linspace(0, 1, n + 1)] * 3
    rng = np.random.default_rng(1)
    b = np.zeros((n + 2,) * 3)
    b[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (n,) * 3) / n ** 3
    prob = W.Problem("dir", v, (W.DIRICHLET,) * 6, b)
    x, st = solve(prob, krylov=bmg.KRYLOV_FCG, precision=bmg.MIXED)
    assert st.final_residual <= 1e-8 * st.initial_residual
    o = oracle.Oracle(v, prob.bc)
    o.assemble()
    o.pin_and_eliminate()
    xe = direct.solve_sparse(o.stencil(0), b, prob.bc)
    assert direct.rel_error(x[1:-1, 1:-1, 1:-1], xe, False) < 1e-6


def _periodic_x(prob):
    """The same source with periodic x faces; the coefficient of the x face
    shared by the first and the last cell through the periodic boundary is
    stored twice (lower face of cell 1, upper face of cell N) and is given
    one value, so that the operator stays symmetric."""
    bc = (W.PERIODIC, W.PERIODIC) + tuple(prob.bc[2:])
    lam = [a.copy() for a in prob.lam]
    n = prob.n[0]
    shared = 0.5 * (lam[0][1] + lam[0][n + 1])
    lam[0][1] = shared
    lam[0][n + 1] = shared
    return W.Problem(prob.name + "_px", prob.vertices, bc, prob.b, lam=lam)


@pytest.mark.parametrize("precision", [bmg.FP64, bmg.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", ["spheres32", "spheres32_px"])
def test_deflated_fcg_matches_direct(gpu, case, precision):
    """Deflation by the high-coefficient sets (the spheres, density ratio
    1000): the stopping test is met, the result matches the FP64 sparse
    direct solve, and it takes fewer iterations than FCG without it."""
    prob = W.spheres(32, seed=9001)
    if case.endswith("_px"):
        prob = _periodic_x(prob)
    print(f"[deflation] case={case} precision={precision}", flush=True)
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(BASE, rel_tol=1e-11, precision=precision,
                             krylov=bmg.KRYLOV_FCG, deflation=1))
    ctx.set_laplacian(prob.lam)
    m = ctx.deflation_vectors()
    x, st = ctx.solve(prob.b.copy())
    ctx.close()
    assert 8 <= m <= 32, m
    assert st.final_residual <= 1e-11 * st.initial_residual
    xe = exact(prob)
    err = direct.rel_error(x[1:-1, 1:-1, 1:-1], xe, True)
    _, st0 = solve(prob, rel_tol=1e-11, precision=precision,
                   krylov=bmg.KRYLOV_FCG)
    print(f"[deflation] {case}: {m} vectors, {st.iters} iterations (without: "
          f"{st0.iters}), error vs direct {err:.2e}", flush=True)
    assert err < 1e-6, err
    assert st.iters < st0.iters


def test_deflation_without_contrast_is_plain_fcg(gpu):
    """No coefficient contrast (the channel): no vectors, and the solve is
    bit for bit the one without deflation."""
    prob = W.channel(32, seed=3)
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(BASE, krylov=bmg.KRYLOV_FCG, deflation=1))
    ctx.set_laplacian(prob.lam)
    assert ctx.deflation_vectors() == 0
    x, st = ctx.solve(prob.b.copy())
    ctx.close()
    x0, st0 = solve(prob, krylov=bmg.KRYLOV_FCG)
    assert st.iters == st0.iters
    assert np.array_equal(x.view(np.int64), x0.view(np.int64))


@pytest.mark.parametrize("krylov", [bmg.KRYLOV_NONE, bmg.KRYLOV_FCG],
                         ids=["mg", "fcg"])
@pytest.mark.parametrize("case", ["spheres32", "channel32"])
def test_mixed_solve_after_nonfinite_fields(gpu, case, krylov):
    """The FP32 cycles start from x = 0 without zeroing it where the first
    red-black sweep rewrites every cell; ghosts behind eliminated faces keep
    what earlier cycles left and enter only multiplied by 0.  A solve start
    zeroes the FP32 fields, so an iteration on a residual that is not a
    number (which leaves NaN in every FP32 field, the coarse ghosts through
    their boundary update) does not reach the next solve: it repeats the
    first solve bit for bit."""
    prob = {"spheres32": lambda: W.spheres(32, seed=6, nspheres=8),
            "channel32": lambda: W.channel(32, seed=6)}[case]()
    print(f"[stale fields] case={case} krylov={krylov}", flush=True)
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(BASE, precision=bmg.MIXED, krylov=krylov))
    ctx.set_laplacian(prob.lam)
    x1, st1 = ctx.solve(prob.b.copy())
    ctx.upload(bmg.R, 0, np.full(ctx.padded(0), np.nan))
    ctx.time_op(bmg.TIME_ITERATION, precision=bmg.MIXED, reps=1)
    assert np.isnan(ctx.download(bmg.X, 0)).any()
    x2, st2 = ctx.solve(prob.b.copy())
    ctx.close()
    assert st1.final_residual <= 1e-8 * st1.initial_residual
    assert st2.iters == st1.iters, (st1.iters, st2.iters)
    assert np.array_equal(x1.view(np.int64), x2.view(np.int64))
