"""The CPU restatement of the reference FFT Poisson solver
(oracle/fft_ref.py) pinned without the reference binary: (1) manufactured
solutions -- a random field with boundary ghosts set from random boundary
values, its finite-difference Laplacian as the source -- are recovered to
rounding for every boundary pair on every axis, on uniform and stretched
meshes; (2) on the channel mesh its solution equals the direct FFT+Thomas
solve of the finite-volume system the multigrid solves (oracle/direct.py)
scaled by the cell volumes; (3) the eigenvalues are those of the 1-D
second difference with each boundary pair (dense eigendecomposition)."""
import itertools

import numpy as np
import pytest

from oracle import direct, fft_ref as F
from briscola_amd import workloads as W


def set_ghosts(x, bc, vertices, vals):
    n = x.shape
    for face in range(6):
        d, side = divmod(face, 2)
        g = [slice(None)] * 3
        c = [slice(None)] * 3
        N = n[d] - 2
        g[d] = 0 if side == 0 else N + 1
        if bc[face] == F.PER:
            c[d] = N if side == 0 else 1
            x[tuple(g)] = x[tuple(c)]
            continue
        c[d] = 1 if side == 0 else N
        h = np.diff(vertices[d])
        hh = h[0] if side == 0 else h[-1]
        if bc[face] == F.DIR:
            x[tuple(g)] = 2 * vals[face] - x[tuple(c)]
        else:
            x[tuple(g)] = x[tuple(c)] + vals[face] * hh
    return x


def manufactured(n, bc, vertices, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros(tuple(q + 2 for q in n))
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal(n)
    x = set_ghosts(x, bc, vertices, rng.standard_normal(6))
    b = np.zeros_like(x)
    b[1:-1, 1:-1, 1:-1] = -F.operator(x, vertices, bc)
    return x, b


def check(x, sol, bc):
    xi = x[1:-1, 1:-1, 1:-1]
    if F.DIR not in bc:
        return np.abs((sol - sol.mean()) - (xi - xi.mean())).max()
    return np.abs(sol - xi).max()


@pytest.mark.parametrize("stretch", [None, 0, 1, 2])
def test_manufactured_all_pairs(stretch):
    n = (12, 16, 10)
    worst = 0.0
    for kinds in itertools.product([F.NEU, F.DIR, F.PER], repeat=3):
        verts = [np.linspace(0.0, 1.0 + d, q + 1) for d, q in enumerate(n)]
        if stretch is not None:
            t = np.linspace(-1.0, 1.0, n[stretch] + 1)
            verts[stretch] = np.tanh(1.5 * t) / np.tanh(1.5)
        bc = [k for k in kinds for _ in (0, 1)]
        x, b = manufactured(n, bc, verts, seed=len(bc) + sum(kinds))
        worst = max(worst, check(x, F.solve(b, x, bc, verts), bc))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("pair", [(F.DIR, F.NEU), (F.NEU, F.DIR)])
def test_manufactured_mixed_pairs(axis, pair):
    n = (8, 12, 16)
    verts = [np.linspace(0.0, 1.0, q + 1) for q in n]
    bc = [F.PER] * 6
    bc[2 * axis], bc[2 * axis + 1] = pair
    x, b = manufactured(n, bc, verts, seed=axis)
    assert check(x, F.solve(b, x, bc, verts), bc) < 1e-12


def test_channel_equals_finite_volume_solve():
    p = W.channel(16, seed=3)
    xv = direct.solve_fft_thomas(p.b, p.vertices, bc_y=(0, 0))
    dx, dy, dz = (np.diff(v) for v in p.vertices)
    V = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    b = np.zeros_like(p.b)
    b[1:-1, 1:-1, 1:-1] = -p.b[1:-1, 1:-1, 1:-1] / V
    xs = F.solve(b, np.zeros_like(p.b), p.bc, p.vertices)
    err = np.abs((xs - xs.mean()) - (xv - xv.mean())).max() / np.abs(xv).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("code,lo,hi", [(1, F.DIR, F.DIR), (2, F.NEU, F.NEU),
                                        (3, F.DIR, F.NEU), (4, F.NEU, F.DIR),
                                        (5, F.PER, F.PER)])
def test_eigenvalues_of_the_second_difference(code, lo, hi):
    n = 9
    T = np.zeros((n, n))
    for i in range(n):
        T[i, i] = -2.0
        if i > 0:
            T[i, i - 1] = 1.0
        if i < n - 1:
            T[i, i + 1] = 1.0
    if lo == F.PER:
        T[0, -1] = T[-1, 0] = 1.0
    else:
        T[0, 0] += -1.0 if lo == F.DIR else 1.0
        T[-1, -1] += -1.0 if hi == F.DIR else 1.0
    assert F.bc_code(lo, hi) == code
    ev = np.sort(np.linalg.eigvals(T).real)
    assert np.allclose(np.sort(F.eigenvalues(code, n)), ev, atol=1e-12)


def test_plan_picks_the_non_uniform_axis():
    u = np.linspace(0, 1, 9)
    s = np.tanh(np.linspace(-1, 1, 9))
    assert F.plan([u, s, u])[0] == 1
    assert F.plan([u, u, s])[0] == 2
    assert F.plan([u, u, u]) == (0, 2, 1)
    with pytest.raises(ValueError):
        F.plan([s, s, u])


def _pinned(bc, verts):
    """True when the reference pins the zero mode: both transformed axes
    Neumann-Neumann or periodic and the solve axis Neumann-Neumann or
    periodic."""
    s, d1, d2 = F.plan(verts)
    codes = [F.bc_code(bc[2 * d], bc[2 * d + 1]) for d in range(3)]
    return codes[s] in (2, 5) and codes[d1] in (2, 5) and codes[d2] in (2, 5)


def manufactured_ddt(n, bc, vertices, dT, seed):
    """x with boundary ghosts and x_old random; b chosen so that x solves
    (laplacian - 1/dT) x = -b - x_old/dT."""
    x, _ = manufactured(n, bc, vertices, seed)
    rng = np.random.default_rng(seed + 100)
    x_old = np.zeros_like(x)
    x_old[1:-1, 1:-1, 1:-1] = rng.standard_normal(n)
    b = np.zeros_like(x)
    xi = x[1:-1, 1:-1, 1:-1]
    b[1:-1, 1:-1, 1:-1] = (-(F.operator(x, vertices, bc) - xi / dT)
                           - x_old[1:-1, 1:-1, 1:-1] / dT)
    return x, b, x_old


@pytest.mark.parametrize("stretch", [None, 1])
def test_manufactured_ddt_all_pairs(stretch):
    """With the time term (FFTPoissonSolver.C:116-131, tridiagonalSolver.C:
    352, :430) every pair on every axis without a pinned zero mode recovers
    the manufactured field exactly, no constant left free; a dropped
    1/dT on the diagonal or in the source fails it."""
    n = (12, 16, 10)
    dt, frac = 2e-3, 0.5
    worst, cases = 0.0, 0
    for kinds in itertools.product([F.NEU, F.DIR, F.PER], repeat=3):
        verts = [np.linspace(0.0, 1.0 + d, q + 1) for d, q in enumerate(n)]
        if stretch is not None:
            t = np.linspace(-1.0, 1.0, n[stretch] + 1)
            verts[stretch] = np.tanh(1.5 * t) / np.tanh(1.5)
        bc = [k for k in kinds for _ in (0, 1)]
        if _pinned(bc, verts):
            continue
        x, b, x_old = manufactured_ddt(n, bc, verts, dt * frac, seed=sum(kinds))
        sol = F.solve(b, x, bc, verts, ddt=(dt, frac), x_old=x_old)
        worst = max(worst, np.abs(sol - x[1:-1, 1:-1, 1:-1]).max())
        cases += 1
    assert cases == 19
    assert worst < 1e-10, worst


@pytest.mark.parametrize("bc", [(F.NEU,) * 6, (F.PER,) * 6,
                                (F.NEU, F.NEU, F.PER, F.PER, F.NEU, F.NEU)])
def test_ddt_pinned_zero_mode_changes_only_the_plane_means(bc):
    """Meshes whose zero mode the reference pins: the pinned row stays in
    the system with the time term, so only the zero transverse mode (the
    mean over each plane of the transformed axes) departs from the
    manufactured field; every other mode is recovered exactly."""
    n = (8, 12, 10)
    verts = [np.linspace(0.0, 1.0, q + 1) for q in n]
    dt, frac = 1e-2, 1.0
    x, b, x_old = manufactured_ddt(n, list(bc), verts, dt * frac, seed=3)
    sol = F.solve(b, x, list(bc), verts, ddt=(dt, frac), x_old=x_old)
    s, d1, d2 = F.plan(verts)
    diff = sol - x[1:-1, 1:-1, 1:-1]
    means = diff.mean(axis=(d1, d2), keepdims=True)
    assert np.abs(diff - means).max() < 1e-10
    assert np.abs(means).max() > 1e-8
