"""The CPU restatement of splitPoissonSolver's modified source
(oracle/split_ref.py) pinned by what the mathematics fixes: no correction
when lambda is lambda0 everywhere, the divergence theorem (the volume sum
of div(phi) telescopes to the boundary faces), and the exact second
difference of a quadratic field on a uniform mesh."""
import numpy as np

from oracle import split_ref as S


def geometry(n, L=(1.0, 1.0, 1.0), stretch=None):
    widths, deltas, centres = [], [], []
    for d in range(3):
        v = np.linspace(0.0, L[d], n[d] + 1)
        if stretch == d:
            t = np.linspace(-1.0, 1.0, n[d] + 1)
            v = L[d] * (np.tanh(1.5 * t) / np.tanh(1.5) + 1) / 2
        h = np.diff(v)
        w = np.concatenate([[h[0]], h, [h[-1]]])
        cc = np.concatenate([[v[0] - h[0] / 2], (v[1:] + v[:-1]) / 2,
                             [v[-1] + h[-1] / 2]])
        dl = np.zeros(n[d] + 2)
        dl[1:] = 1.0 / np.abs(cc[1:] - cc[:-1])
        widths.append(w)
        deltas.append(dl)
        centres.append(cc)
    return widths, deltas, centres


def test_no_correction_when_lambda_is_uniform():
    n = (6, 5, 7)
    w, dl, _ = geometry(n)
    rng = np.random.default_rng(1)
    shp = tuple(q + 2 for q in n)
    x = rng.standard_normal(shp)
    b = rng.standard_normal(shp)
    lam = [np.full(shp, 2.5) for _ in range(3)]
    bh, ph = S.source(x, lam, 2.5, b, w, dl)
    assert all(not p.any() for p in ph)
    assert np.array_equal(bh[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1] / 2.5)


def test_divergence_telescopes_to_the_boundary():
    n = (6, 5, 7)
    w, dl, _ = geometry(n, stretch=1)
    rng = np.random.default_rng(2)
    shp = tuple(q + 2 for q in n)
    x = rng.standard_normal(shp)
    lam = [rng.uniform(0.1, 1.0, shp) for _ in range(3)]
    bh, ph = S.source(x, lam, 1.0, None, w, dl)
    V = (w[0][1:-1, None, None] * w[1][None, 1:-1, None]) * w[2][None, None, 1:-1]
    total = (V * bh[1:-1, 1:-1, 1:-1]).sum()
    bnd = 0.0
    for d in range(3):
        lo = [slice(1, -1)] * 3
        hi = [slice(1, -1)] * 3
        lo[d] = 1
        hi[d] = n[d] + 1
        bnd += ph[d][tuple(lo)].sum() - ph[d][tuple(hi)].sum()
    assert abs(total - bnd) <= 1e-12 * np.abs(np.concatenate([p.ravel() for p in ph])).sum()


def test_quadratic_field_second_difference():
    """lambda/lambda0 - 1 = c on every face and x = sum_d a_d x_d^2: the
    flux divergence is c * sum_d 2 a_d exactly in every interior cell."""
    n = (8, 6, 10)
    w, dl, cc = geometry(n, L=(1.0, 2.0, 0.5))
    a = (0.7, -1.3, 2.1)
    x = (a[0] * cc[0][:, None, None] ** 2 + a[1] * cc[1][None, :, None] ** 2
         + a[2] * cc[2][None, None, :] ** 2)
    shp = x.shape
    c = -0.4
    lam0 = 2.0
    lam = [np.full(shp, lam0 * (1 + c)) for _ in range(3)]
    bh, _ = S.source(x, lam, lam0, None, w, dl)
    exact = c * 2 * sum(a)
    assert np.abs(bh[1:-1, 1:-1, 1:-1] - exact).max() < 1e-9 * abs(exact) * 100
