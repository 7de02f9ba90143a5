"""CPU restatement of the reference FFT Poisson solver, for one process.

Test infrastructure only: the GPU FFT solver (briscola_amd/csrc/bmg_fft.hip)
is checked against it, never built on it.

The reference solves div(grad x) = -b in finite-difference form on a
rectilinear mesh with at least two uniform directions
(src/briscolaFiniteVolume/solvers/PoissonSolvers/FFTPoissonSolver/):

* planner.C:76-183 -- the tridiagonal ("solve") direction is the first
  non-uniform axis, else x; the two others are transformed;
* FourierTransforms.C:133-215 -- per axis the boundary pair becomes a code:
  5 periodic, 1 Dirichlet-Dirichlet, 2 Neumann-Neumann, 3 Dirichlet-Neumann,
  4 Neumann-Dirichlet; FourierTransforms.C:224-254 -- the transform of each
  code (FFTW R2HC/HC2R, RODFT10/01, REDFT10/01, RODFT11, REDFT11), the
  backward transforms unnormalised, the result divided by N (periodic) or
  2N per transformed axis;
* FFTPoissonSolver.C:97-180 -- right-hand side -b, minus the boundary terms
  (x_ghost + x_P)/h^2 (Dirichlet) or (x_ghost - x_P)/h^2 (Neumann) of every
  boundary cell, h the boundary cell's width, from the ghost values x holds;
* tridiagonalSolver.C:14-105 -- eigenvalues of the transformed directions;
  :107-265 -- main diagonal lambda1/h1^2 + lambda2/h2^2 - 1/(d3e h) -
  1/(h d3w) with the boundary rows of each code and the pinned zero mode,
  upper 1/(h d3e), lower 1/(h d3w), d3e/d3w the means of neighbouring
  widths; :313-377 -- Thomas elimination in this order; :379-452 -- the
  cyclic (periodic solve direction) variant by two auxiliary solves.

The scipy transforms are the FFTW ones (DCT/DST types 2, 3 and 4 without
normalisation); the periodic axes use the complex FFT, which pairs every
frequency with the same eigenvalue as the half-complex layout.
"""
import numpy as np
import scipy.fft as sf

NEU, DIR, PER = 0, 1, 2


def bc_code(lo, hi):
    if lo == PER:
        return 5
    if lo == DIR:
        return 1 if hi == DIR else 3
    return 4 if hi == DIR else 2


def is_uniform(v, rtol=1e-12):
    h = np.diff(np.asarray(v, np.float64))
    return bool(np.all(np.abs(h - h[0]) <= rtol * abs(h[0])))


def plan(vertices):
    """(solve axis, first transformed axis, second transformed axis)."""
    n = [len(v) - 1 for v in vertices]
    nonu = [d for d in range(3) if not is_uniform(vertices[d])]
    if len(nonu) > 1:
        raise ValueError("at least two mesh directions must be uniform")
    s = nonu[0] if nonu else (0 if n[0] > 1 else 2)
    d1, d2 = {0: (2, 1), 1: (0, 2), 2: (0, 1)}[s]
    return s, d1, d2


def eigenvalues(code, n):
    i = np.arange(n, dtype=np.float64)
    pi = np.pi
    if code == 1:
        return -4.0 * np.sin((i + 1) * pi / (2.0 * n)) ** 2
    if code == 2:
        return -4.0 * np.sin(i * pi / (2.0 * n)) ** 2
    if code in (3, 4):
        return -4.0 * np.sin((2.0 * (i + 1) - 1.0) * pi / (4.0 * n)) ** 2
    return -4.0 * np.sin(i * pi / n) ** 2


def _fwd(a, code, axis):
    if code == 5:
        return np.fft.fft(a, axis=axis)
    kind = {1: ("dst", 2), 2: ("dct", 2), 3: ("dst", 4), 4: ("dct", 4)}[code]
    f = sf.dst if kind[0] == "dst" else sf.dct
    return f(a, type=kind[1], axis=axis)


def _bwd(a, code, axis):
    """Inverse transform including the reference's normalisation."""
    n = a.shape[axis]
    if code == 5:
        return np.fft.ifft(a, axis=axis)
    kind = {1: ("dst", 3), 2: ("dct", 3), 3: ("dst", 4), 4: ("dct", 4)}[code]
    f = sf.dst if kind[0] == "dst" else sf.dct
    return f(a, type=kind[1], axis=axis) / (2.0 * n)


def _widths_with_ghosts(v, periodic):
    h = np.diff(np.asarray(v, np.float64))
    lo, hi = (h[-1], h[0]) if periodic else (h[0], h[-1])
    return np.concatenate([[lo], h, [hi]])   # index k+1 holds h_k


def rhs(b, x, bc, vertices, x_old=None, dT=None):
    """-b (minus x_old/dT with the time term, FFTPoissonSolver.C:116-131)
    minus the boundary terms, interior cells (FFTPoissonSolver.C:133-180),
    faces in order -x,+x,-y,+y,-z,+z."""
    f = -np.asarray(b, np.float64)[1:-1, 1:-1, 1:-1].copy()
    if x_old is not None:
        f = f - np.asarray(x_old, np.float64)[1:-1, 1:-1, 1:-1] / dT
    n = f.shape
    for face in range(6):
        d, side = divmod(face, 2)
        if bc[face] == PER:
            continue
        h = np.diff(np.asarray(vertices[d], np.float64))
        hh = h[0] if side == 0 else h[-1]
        sq = hh * hh
        inner = [slice(1, q + 1) for q in n]
        ghost = list(inner)
        cell = list(inner)
        ghost[d] = slice(0, 1) if side == 0 else slice(n[d] + 1, n[d] + 2)
        cell[d] = slice(1, 2) if side == 0 else slice(n[d], n[d] + 1)
        out = [slice(None)] * 3
        out[d] = slice(0, 1) if side == 0 else slice(n[d] - 1, n[d])
        xg = x[tuple(ghost)]
        xc = x[tuple(cell)]
        term = (xg + xc) / sq if bc[face] == DIR else (xg - xc) / sq
        f[tuple(out)] -= term
    return f


def solve(b, x, bc, vertices, ddt=None, x_old=None):
    """Interior solution of the reference FFT solver; b and x ghost-padded
    (x supplies the boundary ghost values).  ddt = (delta_t, dt_frac) with
    x_old adds the implicit time term with dT = delta_t*dt_frac: the source
    minus x_old/dT and every diagonal, the pinned zero-mode row included,
    minus 1/dT (tridiagonalSolver.C:334-352, :430)."""
    s, d1, d2 = plan(vertices)
    n = [len(v) - 1 for v in vertices]
    codes = [bc_code(bc[2 * d], bc[2 * d + 1]) for d in range(3)]
    dT = ddt[0] * ddt[1] if ddt is not None else None
    f = rhs(b, x, bc, vertices, x_old if ddt is not None else None,
            dT).astype(np.complex128)
    f = _fwd(f, codes[d1], d1)
    f = _fwd(f, codes[d2], d2)
    lam1 = eigenvalues(codes[d1], n[d1])
    lam2 = eigenvalues(codes[d2], n[d2])
    h1 = np.diff(np.asarray(vertices[d1], np.float64))[0]
    h2 = np.diff(np.asarray(vertices[d2], np.float64))[0]
    d1sqr, d2sqr = h1 * h1, h2 * h2
    hg = _widths_with_ghosts(vertices[s], codes[s] == 5)
    ns = n[s]
    d3 = hg[1:-1]
    d3e = (hg[1:-1] + hg[2:]) / 2.0
    d3w = (hg[1:-1] + hg[:-2]) / 2.0
    DU = 1.0 / (d3 * d3e)
    DL = 1.0 / (d3 * d3w)
    # modes as (n[d1], n[d2], ns) with the solve axis last
    F = np.moveaxis(f, (d1, d2, s), (0, 1, 2))
    L = (lam1[:, None] / d1sqr + lam2[None, :] / d2sqr)
    D = np.empty(F.shape, np.float64)
    for k in range(1, ns - 1):
        D[:, :, k] = L - 1.0 / (d3e[k] * d3[k]) - 1.0 / (d3[k] * d3w[k])
    code = codes[s]
    last = ns - 1
    zero = (lam1[:, None] == 0) & (lam2[None, :] == 0)
    pin = -1.0 / (d3e[0] * d3[0]) - 2.0 / (d3[0] * d3w[0])
    if code == 1:
        D[:, :, 0] = L - 1.0 / (d3e[0] * d3[0]) - 2.0 / (d3[0] * d3w[0])
        D[:, :, last] = L - 2.0 / (d3e[last] * d3[last]) - 1.0 / (d3[last] * d3w[last])
    elif code == 2:
        D[:, :, 0] = np.where(zero, pin, L - 1.0 / (d3e[0] * d3[0]))
        D[:, :, last] = L - 1.0 / (d3[last] * d3w[last])
    elif code == 3:
        D[:, :, 0] = L - 1.0 / (d3e[0] * d3[0]) - 2.0 / (d3[0] * d3w[0])
        D[:, :, last] = L - 1.0 / (d3[last] * d3w[last])
    elif code == 4:
        D[:, :, 0] = L - 1.0 / (d3e[0] * d3[0])
        D[:, :, last] = L - 2.0 / (d3e[last] * d3[last]) - 1.0 / (d3[last] * d3w[last])
    else:
        D[:, :, 0] = np.where(zero, pin, L - 1.0 / (d3e[0] * d3[0]) - 1.0 / (d3[0] * d3w[0]))
        D[:, :, last] = L - 1.0 / (d3e[last] * d3[last]) - 1.0 / (d3[last] * d3w[last])
    if ddt is not None:
        D = D - 1.0 / dT

    def thomas(fr, start):
        Dh = D.copy()
        fh = fr.copy()
        p = np.zeros_like(fr)
        for k in range(1 + start, ns):
            m = DL[k] / Dh[:, :, k - 1]
            Dh[:, :, k] -= m * DU[k - 1]
            fh[:, :, k] -= m * fh[:, :, k - 1]
        p[:, :, last] = fh[:, :, last] / Dh[:, :, last]
        for k in range(ns - 2, start - 1, -1):
            p[:, :, k] = (fh[:, :, k] - DU[k] * p[:, :, k + 1]) / Dh[:, :, k]
        return p

    if code != 5:
        P = thomas(F, 0)
    else:
        u = thomas(F, 1)
        vf = np.zeros(F.shape, np.float64)
        vf[:, :, 1] = -DL[1]
        vf[:, :, last] = -DU[last]
        v = thomas(vf, 1)
        p0 = ((F[:, :, 0] - DL[0] * u[:, :, last] - DU[0] * u[:, :, 1])
              / (D[:, :, 0] + DL[0] * v[:, :, last] + DU[0] * v[:, :, 1]))
        P = u + p0[:, :, None] * v
        P[:, :, 0] = p0
    out = np.moveaxis(P, (0, 1, 2), (d1, d2, s))
    out = _bwd(out, codes[d2], d2)
    out = _bwd(out, codes[d1], d1)
    return np.real(out)


def operator(x, vertices, bc):
    """The finite-difference Laplacian the solver inverts applied to a
    ghost-padded field whose ghosts hold the boundary values (the second
    difference with the reference's widths: (x_+ - x)/(h d3e) - (x - x_-)/
    (h d3w) per axis, walls through the ghost cells).  Interior result."""
    out = np.zeros(tuple(len(v) - 1 for v in vertices))
    for d in range(3):
        h = _widths_with_ghosts(vertices[d], bc[2 * d] == PER)
        hk = h[1:-1]
        de = (h[1:-1] + h[2:]) / 2.0
        dw = (h[1:-1] + h[:-2]) / 2.0
        shp = [1, 1, 1]
        shp[d] = -1
        sl = [slice(1, -1)] * 3
        up = list(sl)
        dn = list(sl)
        up[d] = slice(2, None)
        dn[d] = slice(0, -2)
        xc = x[tuple(sl)]
        out += ((x[tuple(up)] - xc) / (hk * de).reshape(shp)
                - (xc - x[tuple(dn)]) / (hk * dw).reshape(shp))
    return out
