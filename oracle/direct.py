"""FP64 direct solvers of the assembled finite-volume Poisson systems.

Test infrastructure only (checkers, never the measured path).  They solve the
same eliminated 7-point system the multigrid solves, exactly up to rounding,
so that every multigrid result can state its error against an FP64 oracle:

* ``solve_dct``: uniform mesh, all faces Neumann -- the operator is
  diagonalised by the DCT-II (eigenvalues ``sum_d 2 h_perp (cos(pi k/N)-1)``
  of the eliminated stencil); the null-space mode is dropped.
* ``solve_fft_thomas``: uniform periodic x and z, any rectilinear y with
  Neumann or Dirichlet walls -- FFT in x and z, a tridiagonal (Thomas) solve in
  y per wave-number pair.  This is the algorithm of the reference's own FFT
  Poisson solver (src/briscolaFiniteVolume/solvers/PoissonSolvers/FFT,
  doc/chapters/7_solvers.md:101-197), restated for one process.
* ``solve_sparse``: any operator given as a stencil, scipy sparse LU (small
  sizes).

Inputs/outputs use the padded reference layout; the returned solution has the
same mean-zero normalisation for singular systems.
"""
import numpy as np


def _inner(a):
    return a[1:-1, 1:-1, 1:-1]


def solve_dct(b, h):
    """All-Neumann uniform cube of spacing h (lambda-free operator)."""
    import scipy.fft as sf
    bi = _inner(b)
    n = bi.shape
    lam = [2.0 * (np.cos(np.pi * np.arange(q) / q) - 1.0) for q in n]
    L = h * (lam[0][:, None, None] + lam[1][None, :, None] + lam[2][None, None, :])
    bh = sf.dctn(bi, type=2, norm="ortho")
    L[0, 0, 0] = 1.0
    xh = bh / L
    xh[0, 0, 0] = 0.0
    return sf.idctn(xh, type=2, norm="ortho")


def solve_fft_thomas(b, vertices, bc_y=(0, 0)):
    """Periodic uniform x,z; rectilinear y with Neumann (0) / Dirichlet (1).

    Face coefficient = area * (1/|dcc|) (linearGaussLaplacianScheme.C:86);
    y ghost centres are mirrored, so the wall face delta is the wall cell
    width.  Returns the interior solution (mean removed when singular).
    """
    vx, vy, vz = [np.asarray(v, np.float64) for v in vertices]
    nx, ny, nz = len(vx) - 1, len(vy) - 1, len(vz) - 1
    dx = vx[1] - vx[0]
    dz = vz[1] - vz[0]
    dy = np.diff(vy)
    yc = 0.5 * (vy[:-1] + vy[1:])
    # face deltas in y: interior faces between centres, walls to the mirror
    inv = np.empty(ny + 1)
    inv[1:-1] = 1.0 / np.abs(np.diff(yc))
    inv[0] = 1.0 / dy[0]
    inv[-1] = 1.0 / dy[-1]
    cy = (dx * dz) * inv                       # y-face coefficients
    bi = _inner(b)
    bh = np.fft.fft(np.fft.fft(bi, axis=0), axis=2)
    kx = 2.0 * (np.cos(2.0 * np.pi * np.arange(nx) / nx) - 1.0)
    kz = 2.0 * (np.cos(2.0 * np.pi * np.arange(nz) / nz) - 1.0)
    # diagonal part from x and z second differences: (dy dz / dx) kx + ...
    ex = (dz / dx) * kx                         # multiply by dy_j
    ez = (dx / dz) * kz
    diagxz = ex[:, None] + ez[None, :]          # (nx, nz), times dy_j
    lower = np.zeros(ny)
    upper = np.zeros(ny)
    lower[1:] = cy[1:-1]
    upper[:-1] = cy[1:-1]
    centre = -(cy[:-1] + cy[1:])
    # wall elimination: Neumann C += g, Dirichlet C -= g
    centre[0] += cy[0] if bc_y[0] == 0 else -cy[0]
    centre[-1] += cy[-1] if bc_y[1] == 0 else -cy[-1]
    # Thomas over j for all (kx, kz) at once
    d = centre[None, :, None] + dy[None, :, None] * diagxz[:, None, :]
    rhs = bh.astype(np.complex128)
    singular = bc_y[0] == 0 and bc_y[1] == 0
    if singular:
        # pin the (0,0) mode at j = 0
        d = d.copy()
        rhs = rhs.copy()
    cp = np.zeros_like(d)
    dp = np.zeros_like(rhs)
    dd = d.copy()
    if singular:
        dd[0, 0, 0] = 1.0
    up0 = np.broadcast_to(upper[0], dd[:, 0, :].shape).copy()
    if singular:
        up0[0, 0] = 0.0
        rhs[0, 0, 0] = 0.0
    cp[:, 0, :] = up0 / dd[:, 0, :]
    dp[:, 0, :] = rhs[:, 0, :] / dd[:, 0, :]
    for j in range(1, ny):
        lo = np.broadcast_to(lower[j], dd[:, j, :].shape).copy()
        if singular and j == 1:
            lo[0, 0] = 0.0
        m = dd[:, j, :] - lo * cp[:, j - 1, :]
        cp[:, j, :] = upper[j] / m
        dp[:, j, :] = (rhs[:, j, :] - lo * dp[:, j - 1, :]) / m
    xh = np.zeros_like(rhs)
    xh[:, -1, :] = dp[:, -1, :]
    for j in range(ny - 2, -1, -1):
        xh[:, j, :] = dp[:, j, :] - cp[:, j, :] * xh[:, j + 1, :]
    x = np.real(np.fft.ifft(np.fft.ifft(xh, axis=2), axis=0))
    if singular:
        x -= x.mean()
    return x


def solve_sparse(A, b, bc):
    """scipy sparse LU of an eliminated padded stencil (interior unknowns).

    Singular systems (no Dirichlet face) get cell 0 pinned and the mean
    removed from the answer."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    Ai = _inner(A)
    nx, ny, nz = Ai.shape[:3]
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    offs = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    rows, cols, vals = [], [], []
    dims = (nx, ny, nz)
    for s, off in enumerate(offs):
        v = Ai[..., s]
        ok = v != 0
        c = []
        for d, (arr, o) in enumerate(zip((I, J, K), off)):
            q = arr + o
            out = (q < 0) | (q >= dims[d])
            if bc[2 * d] == 2:
                q = q % dims[d]
            else:
                ok = ok & ~out
            c.append(q)
        rows.append(idx[ok])
        cols.append(idx[c[0][ok] % nx, c[1][ok] % ny, c[2][ok] % nz])
        vals.append(v[ok])
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(idx.size,) * 2)
    rhs = _inner(b).ravel().copy()
    singular = 1 not in bc
    if singular:
        M = M.tolil()
        M[0, :] = 0
        M[0, 0] = 1.0
        rhs[0] = 0.0
        M = M.tocsc()
    x = spla.spsolve(M.tocsc(), rhs).reshape(nx, ny, nz)
    if singular:
        x -= x.mean()
    return x


def rel_error(x, x_exact, singular):
    """Relative L2 error, mean removed on both when the system is singular."""
    xi = _inner(x) if x.ndim == 3 and x.shape != x_exact.shape else x
    a = xi - xi.mean() if singular else xi
    e = x_exact - x_exact.mean() if singular else x_exact
    return float(np.linalg.norm(a - e) / np.linalg.norm(e))
