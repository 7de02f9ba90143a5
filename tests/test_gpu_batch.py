"""Batched solves (bmg_solve_batch): independent systems of different
shapes, boundary conditions and precisions solved together.  Every kernel of
an iteration is one launch over all systems, so each system must still get
exactly what its own solve gives it: the solution, the returned source and
the iteration count and residuals of the batched solve are compared bitwise
with bmg_solve_device of the same system alone (whose FP64 path is itself
bitwise the oracle's, tests/test_gpu_parity.py).  Also: the launch grids and
tile shapes the autotuner picks do not change any result, a second solve replays the
captured iteration, and a system with a zero source drops out at once.
Batched flexible CG is held to the same bar against each system's own FCG
solve."""
import numpy as np
import pytest

from briscola_amd import bmg, workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N, D, P = bmg.NEUMANN, bmg.DIRICHLET, bmg.PERIODIC

# (batch index for the source and faces, shape, faces or None, precision);
# every size halves down to at most 3 through even sizes
SYSTEMS = [
    (0, (16, 16, 16), None, bmg.FP64),
    (1, (24, 24, 24), None, bmg.MIXED),
    (2, (12, 24, 6), None, bmg.FP64),
    (3, (16, 8, 32), None, bmg.MIXED),
    (4, (24, 16, 12), (P, P, N, N, D, D), bmg.FP64),
    (5, (16, 16, 16), (N, N, P, P, N, N), bmg.MIXED),
    (6, (8, 12, 16), (D, D, N, N, P, P), bmg.FP64),
    (7, (32, 32, 32), (P, P, P, P, P, P), bmg.MIXED),
    (8, (12, 12, 12), (N, N, N, N, N, N), bmg.FP64),
    (9, (24, 16, 8), (D, D, D, D, D, D), bmg.MIXED),
]


def problem(i, shape, faces):
    p = W.batch_system(i, shape=shape)
    if faces is not None:
        p.bc = tuple(faces)
        b = p.b.copy()
        if D not in faces:
            inner = b[1:-1, 1:-1, 1:-1]
            inner -= inner.mean()
        p.b = b
    return p


def context(p, precision, krylov=bmg.KRYLOV_NONE):
    ctx = bmg.Context(p.vertices, p.bc, p.bc_value, rel_tol=1e-8,
                      tolerance=0.0, max_iter=200, n_sweeps_pre=2,
                      n_sweeps_post=2, cycle=bmg.VCYCLE,
                      coarse_mode=bmg.COARSE_DIRECT, coarse_level=0,
                      precision=precision, krylov=krylov)
    ctx.set_laplacian(p.lam)
    return ctx


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def solve_alone(systems, krylov=bmg.KRYLOV_NONE):
    out = []
    for i, shape, faces, prec in systems:
        p = problem(i, shape, faces)
        ctx = context(p, prec, krylov)
        b, x = dev(p.b), dev(np.zeros_like(p.b))
        st = ctx.solve_device(b.data_ptr(), x.data_ptr())
        out.append((x.cpu().numpy(), b.cpu().numpy(), st.iters,
                    st.initial_residual, st.final_residual))
        ctx.close()
    return out


def solve_batched(systems, autotune=True, repeat=1, krylov=bmg.KRYLOV_NONE,
                  tile_shape=0):
    probs = [problem(i, shape, faces) for i, shape, faces, _ in systems]
    ctxs = [context(p, s[3], krylov) for p, s in zip(probs, systems)]
    batch = bmg.Batch(ctxs, autotune=autotune, tile_shape=tile_shape)
    runs = []
    for r in range(repeat):
        bs = [dev(p.b) for p in probs]
        xs = [dev(np.zeros_like(p.b)) for p in probs]
        stats = batch.solve_device([b.data_ptr() for b in bs],
                                   [x.data_ptr() for x in xs],
                                   const_matrix=r > 0)
        runs.append([(x.cpu().numpy(), b.cpu().numpy(), st.iters,
                      st.initial_residual, st.final_residual)
                     for x, b, st in zip(xs, bs, stats)])
    info = batch.info()
    batch.close()
    for c in ctxs:
        c.close()
    return runs, info


def same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64),
                                                 b.view(np.int64))


@pytest.fixture(scope="module")
def alone(gpu):
    return solve_alone(SYSTEMS)


def test_batch_equals_individual_solves(gpu, alone):
    runs, (launches, wg) = solve_batched(SYSTEMS, repeat=2)
    # the merged iteration is far fewer launches than the systems' own
    assert 0 < launches and wg > 0
    for run in runs:
        for k, (got, ref) in enumerate(zip(run, alone)):
            assert got[2] == ref[2], (k, got[2], ref[2])
            assert got[3] == ref[3] and got[4] == ref[4], k
            assert same(got[0], ref[0]), f"system {k}: x differs"
            assert same(got[1], ref[1]), f"system {k}: b differs"
            assert ref[4] / ref[3] <= 1e-8


def test_launch_grids_do_not_change_results(gpu, alone):
    runs, _ = solve_batched(SYSTEMS[:6], autotune=False)
    for k, (got, ref) in enumerate(zip(runs[0], alone[:6])):
        assert got[2] == ref[2] and same(got[0], ref[0]), k


@pytest.mark.parametrize("ws", [3, 4, 5, 6])
def test_tile_shapes_do_not_change_results(gpu, alone, ws):
    """Every shape the autotuner may pick (rows of 2^ws cells per wavefront
    row for the phases, residuals, restriction and prolongation), forced on
    every launch, with MG iterations."""
    runs, _ = solve_batched(SYSTEMS, autotune=False, tile_shape=ws)
    for k, (got, ref) in enumerate(zip(runs[0], alone)):
        assert got[2] == ref[2], (ws, k, got[2], ref[2])
        assert same(got[0], ref[0]) and same(got[1], ref[1]), (ws, k)


def test_zero_source_system_drops_out(gpu):
    p0 = problem(0, (16, 16, 16), None)
    p1 = problem(1, (12, 12, 12), (D, D, N, N, N, N))
    p1.b[...] = 0.0
    ctxs = [context(p0, bmg.FP64), context(p1, bmg.FP64)]
    batch = bmg.Batch(ctxs)
    bs = [dev(p0.b), dev(p1.b)]
    xs = [dev(np.zeros_like(p0.b)), dev(np.zeros_like(p1.b))]
    st = batch.solve_device([b.data_ptr() for b in bs],
                            [x.data_ptr() for x in xs])
    assert st[0].iters > 0 and st[1].iters == 0
    assert st[1].initial_residual == 0.0
    assert not xs[1].abs().max().item()
    batch.close()
    for c in ctxs:
        c.close()


# systems for flexible CG: the uneven shapes of SYSTEMS plus anisotropic
# cells (the slow case of plain MG), FP64 and mixed cycles
FCG_SYSTEMS = SYSTEMS[:6] + [
    (10, (16, 8, 32), (N, N, D, D, P, P), bmg.MIXED),
    (11, (32, 16, 64), (D, D, N, N, N, N), bmg.FP64),
]


def test_batch_fcg_equals_individual_solves(gpu):
    """Batched flexible CG (restart p = M r, captured iterations, the
    confirmation on the true residual, all merged over the systems that
    need each) gives every system exactly its own FCG solve."""
    alone = solve_alone(FCG_SYSTEMS, bmg.KRYLOV_FCG)
    runs, (launches, wg) = solve_batched(FCG_SYSTEMS, repeat=2,
                                         krylov=bmg.KRYLOV_FCG)
    assert launches > 0 and wg > 0
    for run in runs:
        for k, (got, ref) in enumerate(zip(run, alone)):
            assert got[2] == ref[2], (k, got[2], ref[2])
            assert got[3] == ref[3] and got[4] == ref[4], k
            assert same(got[0], ref[0]), f"system {k}: x differs"
            assert same(got[1], ref[1]), f"system {k}: b differs"
            assert ref[4] / ref[3] <= 1e-8


def test_batch_rejects_learned_and_mixed_outer_iterations(gpu):
    p = problem(0, (8, 8, 8), None)
    for kinds in ((bmg.KRYLOV_FCG_LEARNED,),
                  (bmg.KRYLOV_NONE, bmg.KRYLOV_FCG)):
        ctxs = []
        for k in kinds:
            ctx = bmg.Context(p.vertices, p.bc, p.bc_value, rel_tol=1e-8,
                              tolerance=0.0, krylov=k)
            ctx.set_laplacian(None)
            ctxs.append(ctx)
        batch = bmg.Batch(ctxs)
        bs = [dev(p.b) for _ in ctxs]
        xs = [dev(np.zeros_like(p.b)) for _ in ctxs]
        with pytest.raises(bmg.BmgError):
            batch.solve_device([b.data_ptr() for b in bs],
                               [x.data_ptr() for x in xs])
        batch.close()
        for c in ctxs:
            c.close()


def _oracle(p, rel_tol):
    import oracle
    o = oracle.Oracle(p.vertices, p.bc, p.bc_value)
    o.set_opts(rel_tol=rel_tol, tolerance=0.0, max_iter=400, n_sweeps_pre=2,
               n_sweeps_post=2, cycle=oracle.VCYCLE,
               coarse_mode=oracle.COARSE_DIRECT, coarse_level=0)
    o.assemble(p.lam)
    o.set_field(oracle.B, 0, p.b)
    it, r0, r1 = o.solve()
    return o, it, r0, r1


def _fp64_reference(p):
    """Sparse LU of the eliminated system (oracle/direct.py), or for the
    fully periodic 32^3 system (whose LU fills in) the oracle's FP64 MG
    to 1e-13."""
    import oracle
    from oracle import direct
    if p.bc == (P,) * 6 and min(p.n) >= 32:
        o, _, _, r1 = _oracle(p, 1e-13)
        assert r1 <= 1e-13
        return o.field(oracle.X, 0)[1:-1, 1:-1, 1:-1]
    o = oracle.Oracle(p.vertices, p.bc, p.bc_value)
    o.assemble(p.lam)
    o.pin_and_eliminate()
    return direct.solve_sparse(o.stencil(0), p.b, p.bc)


def test_batch_against_the_oracle(gpu):
    """The batched solve against the CPU oracle, system by system: the FP64
    systems bitwise equal to the oracle's MG::solve (x, returned b,
    iteration count); the mixed-precision systems, and every system of the
    batched flexible CG, within 1e-7 of an FP64 direct solve."""
    import oracle
    from oracle import direct
    runs, _ = solve_batched(SYSTEMS)
    fruns, _ = solve_batched(FCG_SYSTEMS, krylov=bmg.KRYLOV_FCG)
    for tag, systems, run in (("mg", SYSTEMS, runs[0]),
                              ("fcg", FCG_SYSTEMS, fruns[0])):
        for k, (i, shape, faces, prec) in enumerate(systems):
            p = problem(i, shape, faces)
            got = run[k]
            print(f"[batch vs oracle] {tag} system {k} {shape} {p.bc} "
                  f"prec={prec}", flush=True)
            if tag == "mg" and prec == bmg.FP64:
                o, it, r0, r1 = _oracle(p, 1e-8)
                assert got[2] == it, (k, got[2], it)
                assert np.array_equal(got[0], o.field(oracle.X, 0)), k
                assert np.array_equal(got[1], o.field(oracle.B, 0)), k
            xe = _fp64_reference(p)
            err = direct.rel_error(got[0][1:-1, 1:-1, 1:-1], xe,
                                   D not in p.bc)
            print(f"[batch vs oracle] {tag} system {k}: {got[2]} iterations,"
                  f" error vs FP64 direct {err:.2e}", flush=True)
            assert err < 1e-7, (tag, k, err)
