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
