"""Red-black smoothing on levels with periodic faces, wrapped form.

By default (bmg_set_periodic_smoothing mode 1) each red-black phase of an FP32
cycle covers the whole interior and writes the periodic face ghosts together
with the cells they mirror, on every level whose cell count is even along
each periodic axis; mode 2 does the same in the FP64 cycles; mode 0 keeps the
reference's order everywhere (interior phases, ghost fill, boundary slabs,
rbgsSmoother.C:186-374).  The wrapped form has no reference counterpart, so
it is pinned by:

* a numpy restatement of periodic red-black Gauss-Seidel (same colour
  order, same per-cell arithmetic and rounding order), against which the
  FP64 kernels of mode 2 are bitwise; levels with an odd periodic count keep
  the reference's order, bitwise against the oracle;
* a V-cycle composed from single level operations, bitwise equal to the
  library's own cycle, which zeroes x and runs the first phase in one launch
  (its periodic ghosts included) and skips the ghost refresh of the residual
  on wrapped levels;
* FP32 solves in both modes: they reach the tolerance, match an FP64 direct
  solve (FFT in the periodic axes, Thomas in the wall-normal one), and the
  wrapped phases need at most about as many cycles as the slab order.
"""
import numpy as np
import pytest

import oracle
from oracle import direct
from briscola_amd import bmg, workloads as W
from periodic_rbgs import wrapped_rbgs

pytestmark = pytest.mark.gpu

BASE = dict(rel_tol=1e-8, tolerance=0.0, max_iter=400, n_sweeps_pre=2,
            n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_mode=bmg.COARSE_DIRECT,
            coarse_level=1)


def y_periodic(n=16, seed=0):
    """Dirichlet x faces, periodic y, Neumann z on a uniform mesh."""
    verts = [np.arange(q + 1) / q for q in n]
    rng = np.random.default_rng(seed)
    b = np.zeros(tuple(q + 2 for q in n))
    b[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, n) / np.prod(n)
    return W.Problem("yper", verts,
                     (W.DIRICHLET, W.DIRICHLET, W.PERIODIC, W.PERIODIC,
                      W.NEUMANN, W.NEUMANN), b)


LEVEL_CASES = {
    "channel16": lambda: W.channel(16, seed=1),
    # z counts 6 and 3 below the finest level: reference order there
    "channel_thin": lambda: W.channel((24, 16, 12), seed=2),
    # two cells along z on every level: reference order everywhere
    "thin_z": lambda: W.channel((16, 8, 2), seed=3),
    "y_periodic": lambda: y_periodic((12, 16, 24), seed=4),
    "channel_longz": lambda: W.channel((8, 16, 128), seed=5),
}


def wrapped_level(shape, axes):
    """Levels the library smooths with wrapped phases: an even count of at
    least 8 cells along every periodic axis."""
    return all(shape[d] % 2 == 0 and shape[d] >= 8 for d in axes)


def periodic_axes(prob):
    return [d for d in range(3) if prob.bc[2 * d] == W.PERIODIC]


def assert_same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: max |diff| {d.max():.3e} at "
                             f"{np.unravel_index(np.argmax(d), d.shape)}")


def make(prob, mode, native=True, **opts):
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(BASE, **opts))
    ctx.set_periodic_smoothing(mode)
    if native:
        ctx.set_laplacian(prob.lam)
    return ctx


@pytest.mark.parametrize("native", [False, True], ids=["stencil", "native"])
@pytest.mark.parametrize("case", list(LEVEL_CASES))
def test_wrapped_smoothing_bitwise(gpu, case, native):
    prob = LEVEL_CASES[case]()
    print(f"[wrapped smoothing] case={case} native={native}", flush=True)
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    o.set_opts(**{k: v for k, v in BASE.items()})
    o.assemble(prob.lam)
    ctx = make(prob, 2, native)
    assert ctx.shapes == o.shapes
    if not native:
        for l in range(o.nlevels):
            ctx.set_stencil(l, o.stencil(l))
    o.pin_and_eliminate()
    ctx.prepare()
    axes = periodic_axes(prob)
    rng = np.random.default_rng(11)
    wrapped = 0
    for l in range(o.nlevels - 1):
        shp = o.padded(l)
        x = rng.uniform(-1, 1, shp)
        b = rng.uniform(-1, 1, shp)
        ctx.upload(bmg.X, l, x)
        ctx.upload(bmg.B, l, b)
        ctx.smooth(l, 2)
        got = ctx.download(bmg.X, l)
        if wrapped_level(o.shapes[l], axes):
            want = wrapped_rbgs(o.stencil(l), x, b, axes, 2)
            wrapped += 1
        else:
            o.set_field(oracle.X, l, x)
            o.set_field(oracle.B, l, b)
            o.smooth(l, 2)
            want = o.field(oracle.X, l)
        assert_same(got, want, f"smooth l={l} shape={o.shapes[l]}")
    assert (wrapped > 0) == (case != "thin_z")
    ctx.close()


def composed_vcycle(ctx, nlev, lc, singular):
    """MG::cycle's V-cycle (MG.C:19-129) from single level operations: x of
    the coarser levels zeroed by upload, smoothing from it, the residual of
    levels > 0 with the ghost refresh, the singular system's pinned cell at
    the coarse level, direct coarse solve, prolongation and post-smoothing.
    The level-0 residual is the one present before the cycle (MG.C:72-77)."""
    for l in range(lc):
        if l > 0:
            ctx.upload(bmg.X, l, np.zeros(ctx.padded(l)))
        ctx.smooth(l, BASE["n_sweeps_pre"])
        if l > 0:
            ctx.residual(l, True)
        ctx.restrict(l)
    if singular:
        bc_ = ctx.download(bmg.B, lc)
        bc_[1, 1, 1] = 0.0
        ctx.upload(bmg.B, lc, bc_)
    ctx.upload(bmg.X, lc, np.zeros(ctx.padded(lc)))
    ctx.coarse_solve()
    for l in range(lc - 1, -1, -1):
        ctx.prolong(l)
        ctx.smooth(l, BASE["n_sweeps_post"])


@pytest.mark.parametrize("mode", [0, 2], ids=["reference_order", "wrapped"])
@pytest.mark.parametrize("case", ["channel16", "channel_thin", "y_periodic"])
def test_cycle_equals_composed_level_ops(gpu, case, mode):
    prob = LEVEL_CASES[case]()
    print(f"[composed cycle] case={case} mode={mode}", flush=True)
    x0 = np.random.default_rng(5).uniform(-1, 1, prob.b.shape)
    out = []
    for composed in (False, True):
        ctx = make(prob, mode)
        ctx.prepare()
        ctx.upload(bmg.X, 0, x0)
        ctx.upload(bmg.B, 0, prob.b)
        ctx.residual(0, True)
        nlev = len(ctx.shapes)
        lc = nlev - 1 - BASE["coarse_level"]
        singular = all(f != W.DIRICHLET for f in prob.bc)
        if composed:
            composed_vcycle(ctx, nlev, lc, singular)
        else:
            ctx.cycle()
        out.append(ctx.download(bmg.X, 0))
        ctx.close()
    assert_same(out[0], out[1], "cycle vs composed level operations")


# (cells, what the case exercises)
SOLVE_CASES = {
    "channel32": (32, "even periodic axes on every level"),
    "odd_coarse": ((32, 16, 24), "z counts 6, 3 on the third, fourth level"),
    "thin_levels": ((24, 16, 12), "z counts 6, 3 below the finest level"),
    "thin_z": ((32, 16, 2), "two cells along periodic z: slab order only"),
}


@pytest.mark.parametrize("krylov", [bmg.KRYLOV_NONE, bmg.KRYLOV_FCG],
                         ids=["mg", "fcg"])
@pytest.mark.parametrize("case", list(SOLVE_CASES))
def test_fp32_wrapped_solves_match_direct(gpu, case, krylov):
    prob = W.channel(SOLVE_CASES[case][0], seed=3)
    xe = direct.solve_fft_thomas(prob.b, prob.vertices)
    res = {}
    for mode in (0, 1):
        print(f"[periodic fp32] case={case} krylov={krylov} mode={mode}",
              flush=True)
        ctx = make(prob, mode, precision=bmg.MIXED, krylov=krylov)
        x, st = ctx.solve(prob.b.copy())
        ctx.close()
        assert st.final_residual <= 1e-8 * st.initial_residual, (mode, st)
        err = direct.rel_error(x[1:-1, 1:-1, 1:-1], xe, True)
        assert err < 1e-6, (mode, err)
        res[mode] = st.iters
    print(f"[periodic fp32] {case} ({SOLVE_CASES[case][1]}): iterations "
          f"slab order {res[0]}, wrapped phases {res[1]}", flush=True)
    # another Gauss-Seidel order: the plain cycles of these anisotropic
    # (stretched-wall) systems converge slowly, and their count moves by a
    # few per cent with the order; a stale or misplaced ghost costs more
    assert res[1] <= 1.1 * res[0] + 1, res


def test_periodic_smoothing_mode_is_checked(gpu):
    prob = W.channel(16, seed=1)
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, **BASE)
    with pytest.raises(bmg.BmgError):
        ctx.set_periodic_smoothing(3)
    ctx.close()
