"""Fused red-black smoothing (bmg_set_fused_smoothing): two consecutive
phases in one LDS-staged launch must give exactly the bits of one launch per
phase, which are the oracle's (rbgsSmoother.C:65-183, 376-437).

Mode 2 forces the fused path on every level whose rows fit the kernel's LDS
ring, so small grids exercise partial tiles in J and I, the halo rows and
planes at the domain edges, odd padded z extents (rows whose last position
is not a cell) and the copy of the untouched colour into the spare buffer.
Level smoothing is compared with the oracle over the whole padded field
(ghosts included) for 1, 2 and 3 sweeps: an even-sized interior has three
phases per sweep and an odd phase count (pairs plus a last single phase
that reads the second colour from wherever the last pair left it, then the
copy back), the (16, 12, 3) box four (pairs only).  Whole FP64, mixed and
flexible-CG solves with mode 2 are compared with mode 0 bit for bit, with
the same iteration counts and residuals."""
import numpy as np
import pytest

import oracle
from briscola_amd import bmg, workloads as W

pytestmark = pytest.mark.gpu

N, D = W.NEUMANN, W.DIRICHLET

OPTS = dict(rel_tol=1e-10, tolerance=0.0, n_sweeps_pre=2, n_sweeps_post=2,
            cycle=oracle.VCYCLE, coarse_mode=oracle.COARSE_DIRECT,
            coarse_level=1)


def box(n, bc, seed, stretch=None):
    """Random source on a box with the given (non-periodic) faces; axis
    `stretch` gets a tanh-stretched mesh."""
    verts = [np.arange(q + 1) / q for q in n]
    if stretch is not None:
        t = np.linspace(-1.0, 1.0, n[stretch] + 1)
        verts[stretch] = 0.5 * (1.0 + np.tanh(1.5 * t) / np.tanh(1.5))
    rng = np.random.default_rng(seed)
    b = np.zeros(tuple(q + 2 for q in n))
    b[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, n) / np.prod(n)
    if D not in bc:
        b[1:-1, 1:-1, 1:-1] -= b[1:-1, 1:-1, 1:-1].mean()
    return W.Problem("box", verts, tuple(bc), b)


CASES = {
    "neumann16": lambda: W.uniform_neumann(16, seed=1),
    "spheres24": lambda: W.spheres(24, seed=2, nspheres=5),
    "aniso": lambda: W.uniform_neumann((16, 8, 32), seed=3),
    "dirichlet_mix": lambda: box((24, 12, 16), (D, D, N, N, D, N), 4),
    "stretched": lambda: box((16, 24, 12), (N, N, D, D, N, N), 5, stretch=1),
    # odd padded z extent on level 0 (N_z = 3)
    "thin_z": lambda: box((16, 12, 3), (D, N, N, D, N, N), 6),
}


def context(prob, mode, **extra):
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(OPTS, **extra))
    ctx.set_fused_smoothing(mode)
    return ctx


@pytest.mark.parametrize("sweeps", [1, 2, 3])
@pytest.mark.parametrize("native", [False, True], ids=["stencil", "native"])
@pytest.mark.parametrize("case", list(CASES))
def test_fused_smoothing_matches_oracle(gpu, case, native, sweeps):
    prob = CASES[case]()
    print(f"[fused smooth] case={case} native={native} sweeps={sweeps}",
          flush=True)
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    o.set_opts(**OPTS)
    o.assemble(prob.lam)
    ctx = context(prob, 2)
    if native:
        ctx.set_laplacian(prob.lam)
    else:
        for l in range(o.nlevels):
            ctx.set_stencil(l, o.stencil(l))
    o.pin_and_eliminate()
    o.prepare_coarse()
    ctx.prepare()
    rng = np.random.default_rng(11)
    for l in range(o.nlevels - 1):
        shp = o.padded(l)
        x = rng.uniform(-1, 1, shp)
        b = rng.uniform(-1, 1, shp)
        for which, v in ((oracle.X, x), (oracle.B, b)):
            o.set_field(which, l, v)
            ctx.upload(which, l, v)
        o.smooth(l, sweeps)
        ctx.smooth(l, sweeps)
        got, ref = ctx.download(0, l), o.field(oracle.X, l)
        if not np.array_equal(got, ref):
            d = np.abs(got - ref)
            raise AssertionError(
                f"level {l}: max |diff| {d.max():.3e} at "
                f"{np.unravel_index(np.argmax(d), d.shape)}")
    ctx.close()


@pytest.mark.parametrize("method", ["fp64", "mixed", "fcg_mixed"])
@pytest.mark.parametrize("case", ["neumann16", "spheres24", "dirichlet_mix",
                                  "stretched"])
def test_fused_solve_bitwise(gpu, case, method):
    prob = CASES[case]()
    print(f"[fused solve] case={case} method={method}", flush=True)
    extra = dict(precision=bmg.MIXED if "mixed" in method else bmg.FP64,
                 krylov=bmg.KRYLOV_FCG if method.startswith("fcg")
                 else bmg.KRYLOV_NONE, max_iter=300)
    out = []
    for mode in (0, 2):
        ctx = context(prob, mode, **extra)
        ctx.set_laplacian(prob.lam)
        x, st = ctx.solve(prob.b.copy())
        # a second solve replays the captured iteration
        x2, st2 = ctx.solve(prob.b.copy())
        ctx.close()
        assert np.array_equal(x, x2) and st.iters == st2.iters
        out.append((x, st.iters, st.initial_residual, st.final_residual))
    (x0, i0, a0, f0), (x2, i2, a2, f2) = out
    assert i0 == i2 and a0 == a2 and f0 == f2, (i0, i2, f0, f2)
    assert np.array_equal(x0, x2), np.abs(x0 - x2).max()
    assert f0 / a0 <= 1e-10
