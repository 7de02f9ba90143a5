"""The GPU FFT Poisson solver (bmg_fft_*, briscola_amd/csrc/bmg_fft.hip)
against the CPU restatement of the reference FFT solver (oracle/fft_ref.py),
which is pinned by manufactured solutions (tests/test_fft_oracle.py): every
supported boundary pair on the transformed and tridiagonal axes, uniform and
stretched meshes, inhomogeneous boundary values through the ghosts of x,
and the ghost refresh after the solve.  FFT rounding differs from the
oracle's, so the bar is 1e-11 of the solution's magnitude.  Also: mixed
Dirichlet/Neumann pairs on the transformed axes (quarter-wave transforms),
on the channel the FFT solution equals the multigrid's to its tolerance, and
a mesh with two non-uniform directions is refused.  With the implicit time
term every pair matches the oracle (pinned zero modes included) and
recovers the manufactured field where no zero mode is pinned."""
import itertools

import numpy as np
import pytest

from oracle import fft_ref as F
from briscola_amd import bmg, workloads as W

from test_fft_oracle import manufactured

pytestmark = pytest.mark.gpu

TOL = 1e-11


def rel(a, b, singular):
    if singular:
        a = a - a.mean()
        b = b - b.mean()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def ghost_rule(x, bc, vertices, vals):
    """Expected face ghosts after the solve (periodic image, Neumann
    x_P + g h, Dirichlet 2v - x_P)."""
    from test_fft_oracle import set_ghosts
    y = x.copy()
    return set_ghosts(y, bc, vertices, vals)


KINDS = [F.NEU, F.DIR, F.PER]


@pytest.mark.parametrize("stretch", [None, 1])
@pytest.mark.parametrize("kinds", list(itertools.product(KINDS, repeat=3)),
                         ids=lambda k: "".join("NDP"[q] for q in k))
def test_fft_matches_oracle(gpu, stretch, kinds):
    n = (16, 12, 10)
    verts = [np.linspace(0.0, 1.0 + d, q + 1) for d, q in enumerate(n)]
    if stretch is not None:
        t = np.linspace(-1.0, 1.0, n[stretch] + 1)
        verts[stretch] = np.tanh(1.5 * t) / np.tanh(1.5)
    bc = [k for k in kinds for _ in (0, 1)]
    rng = np.random.default_rng(sum(kinds) + 7)
    vals = [0.0 if k == F.PER else v for k, v in zip(bc, rng.standard_normal(6))]
    x, b = manufactured(n, bc, verts, seed=3)
    from test_fft_oracle import set_ghosts
    x = set_ghosts(x, bc, verts, vals)
    b[1:-1, 1:-1, 1:-1] = -F.operator(x, verts, bc)
    ref = F.solve(b, x, bc, verts)
    print(f"[fft] kinds={kinds} stretch={stretch}", flush=True)
    fft = bmg.FFT(verts, bc, vals)
    xg = x.copy()
    fft.solve(b.copy(), xg)
    singular = F.DIR not in bc
    assert rel(xg[1:-1, 1:-1, 1:-1], ref, singular) < TOL
    # the solution is the manufactured field (up to a constant if singular)
    assert rel(xg[1:-1, 1:-1, 1:-1], x[1:-1, 1:-1, 1:-1], singular) < 1e-10
    # face ghosts follow the boundary conditions of the new solution
    exp = ghost_rule(xg, bc, verts, vals)
    for d in range(3):
        for side in (0, 1):
            sl = [slice(1, -1)] * 3
            sl[d] = 0 if side == 0 else -1
            assert np.allclose(xg[tuple(sl)], exp[tuple(sl)], rtol=0, atol=1e-12)


def test_fft_channel_equals_multigrid(gpu):
    p = W.channel(32, seed=2)
    ctx = bmg.Context(p.vertices, p.bc, rel_tol=1e-12, tolerance=0.0,
                      n_sweeps_pre=2, n_sweeps_post=2, cycle=bmg.VCYCLE,
                      coarse_level=1, krylov=bmg.KRYLOV_FCG)
    ctx.set_laplacian()
    xm, _ = ctx.solve(p.b.copy())
    dx, dy, dz = (np.diff(v) for v in p.vertices)
    V = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    b = np.zeros_like(p.b)
    b[1:-1, 1:-1, 1:-1] = -p.b[1:-1, 1:-1, 1:-1] / V
    fft = bmg.FFT(p.vertices, p.bc)
    xf = fft.solve(b)
    assert rel(xf[1:-1, 1:-1, 1:-1], xm[1:-1, 1:-1, 1:-1], True) < 1e-9


def test_fft_refuses_unsupported(gpu):
    u = np.linspace(0, 1, 9)
    s = np.tanh(np.linspace(-1, 1, 9))
    # two non-uniform directions
    with pytest.raises(bmg.BmgError):
        bmg.FFT([s, s, u], [F.NEU] * 6)


PAIRS_T = [(F.DIR, F.NEU), (F.NEU, F.DIR), (F.NEU, F.NEU), (F.PER, F.PER)]


@pytest.mark.parametrize("solve_pair", [(F.NEU, F.NEU), (F.DIR, F.NEU),
                                        (F.PER, F.PER)],
                         ids=["NN", "DN", "PP"])
@pytest.mark.parametrize("pairs", [q for q in itertools.product(PAIRS_T,
                                                                repeat=2)
                                   if q[0][0] != q[0][1] and q[0][0] != F.PER
                                   or q[1][0] != q[1][1] and q[1][0] != F.PER],
                         ids=lambda q: "-".join("NDP"[a[0]] + "NDP"[a[1]]
                                                for a in q))
@pytest.mark.parametrize("stretch", [None, 1])
def test_fft_quarter_wave_transformed_axes(gpu, stretch, pairs, solve_pair):
    """Dirichlet-Neumann and Neumann-Dirichlet pairs on the transformed
    axes (the reference's DST-IV / DCT-IV, RODFT11 / REDFT11), computed from
    the quarter-wave extension to 4N."""
    n = (12, 16, 10)
    verts = [np.linspace(0.0, 1.0 + d, q + 1) for d, q in enumerate(n)]
    if stretch is not None:
        t = np.linspace(-1.0, 1.0, n[stretch] + 1)
        verts[stretch] = np.tanh(1.5 * t) / np.tanh(1.5)
    s_axis = 0 if stretch is None else stretch
    trans = [d for d in (2, 1, 0) if d != s_axis][:2]
    bc = [None] * 6
    bc[2 * s_axis], bc[2 * s_axis + 1] = solve_pair
    for d, pr in zip(trans, pairs):
        bc[2 * d], bc[2 * d + 1] = pr
    rng = np.random.default_rng(5)
    vals = [0.0 if k == F.PER else v for k, v in zip(bc, rng.standard_normal(6))]
    from test_fft_oracle import set_ghosts
    x, b = manufactured(n, bc, verts, seed=13)
    x = set_ghosts(x, bc, verts, vals)
    b[1:-1, 1:-1, 1:-1] = -F.operator(x, verts, bc)
    ref = F.solve(b, x, bc, verts)
    print(f"[fft quarter-wave] bc={bc} stretch={stretch}", flush=True)
    fft = bmg.FFT(verts, bc, vals)
    xg = fft.solve(b.copy(), x.copy())
    singular = F.DIR not in bc
    assert rel(xg[1:-1, 1:-1, 1:-1], ref, singular) < TOL
    assert rel(xg[1:-1, 1:-1, 1:-1], x[1:-1, 1:-1, 1:-1], singular) < 1e-10


@pytest.mark.parametrize("pair", [(F.DIR, F.NEU), (F.NEU, F.DIR)])
def test_fft_mixed_pair_on_the_tridiagonal_axis(gpu, pair):
    n = (10, 8, 12)
    verts = [np.linspace(0.0, 1.0, q + 1) for q in n]
    bc = [pair[0], pair[1], F.PER, F.PER, F.NEU, F.NEU]
    vals = [0.5, -0.25, 0.0, 0.0, 0.1, 0.2]
    from test_fft_oracle import set_ghosts
    x, b = manufactured(n, bc, verts, seed=11)
    x = set_ghosts(x, bc, verts, vals)
    b[1:-1, 1:-1, 1:-1] = -F.operator(x, verts, bc)
    ref = F.solve(b, x, bc, verts)
    fft = bmg.FFT(verts, bc, vals)
    xg = fft.solve(b.copy(), x.copy())
    assert rel(xg[1:-1, 1:-1, 1:-1], ref, False) < TOL


@pytest.mark.parametrize("stretch", [None, 2])
@pytest.mark.parametrize("kinds", list(itertools.product(KINDS, repeat=3)),
                         ids=lambda k: "".join("NDP"[q] for q in k))
def test_fft_ddt_matches_oracle(gpu, stretch, kinds):
    """(laplacian - 1/dT) x = -b - x_old/dT, dT = deltaT*dtFrac, host and
    device entry points."""
    from test_fft_oracle import manufactured_ddt, _pinned
    torch = pytest.importorskip("torch")
    n = (16, 12, 10)
    verts = [np.linspace(0.0, 1.0 + d, q + 1) for d, q in enumerate(n)]
    if stretch is not None:
        t = np.linspace(-1.0, 1.0, n[stretch] + 1)
        verts[stretch] = np.tanh(1.5 * t) / np.tanh(1.5)
    bc = [k for k in kinds for _ in (0, 1)]
    dt, frac = 4e-3, 0.5
    print(f"[fft ddt] kinds={kinds} stretch={stretch}", flush=True)
    x, b, x_old = manufactured_ddt(n, bc, verts, dt * frac, seed=sum(kinds) + 5)
    ref = F.solve(b, x, bc, verts, ddt=(dt, frac), x_old=x_old)
    fft = bmg.FFT(verts, bc)
    xg = x.copy()
    fft.solve(b.copy(), xg, ddt=(dt, frac), x_old=x_old)
    assert rel(xg[1:-1, 1:-1, 1:-1], ref, False) < TOL
    if not _pinned(bc, verts):
        assert rel(xg[1:-1, 1:-1, 1:-1], x[1:-1, 1:-1, 1:-1], False) < 1e-10
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    db, dx, dold = dev(b), dev(x), dev(x_old)
    fft.solve_device(db.data_ptr(), dx.data_ptr(), ddt=(dt, frac),
                     x_old_ptr=dold.data_ptr())
    assert np.array_equal(dx.cpu().numpy(), xg)
    with pytest.raises(bmg.BmgError):
        fft.solve(b.copy(), x.copy(), ddt=(dt, frac), x_old=None)
    with pytest.raises(bmg.BmgError):
        fft.solve(b.copy(), x.copy(), ddt=(0.0, frac), x_old=x_old)
    fft.close()
