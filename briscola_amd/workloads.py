"""Synthetic Poisson workloads (meshes, boundary kinds, right-hand sides).

Every random draw comes from a SplitMix64 counter stream keyed by
``(case, seed, field)``: draw ``i`` of a stream is ``mix64(base + (i+1)*G)``
with ``G = 0x9E3779B97F4A7C15`` and ``base = mix64(case<<40 ^ seed<<20 ^
field)``, so a draw never depends on how many draws came before it and the
same numbers come out in any language.  Uniforms on [-1,1) are
``(z >> 11) * 2**-52 - 1``.

Cases (sizes are parameters; the defaults are the benchmark shapes):

* ``uniform_neumann``: unit cube, all faces Neumann(0), ``f ~ U[-1,1)`` per
  cell, mean removed, ``b = h^3 f``.
* ``channel``: [0,4pi] x [-1,1] x [0,2pi], periodic in x and z, Neumann in y,
  ``y_j = tanh(beta (2j/N - 1))/tanh(beta)``; ``f`` = sum of cosine modes with
  integer wave numbers ``0 < |m| <= 8`` and amplitude ``|m|^(-5/3)`` (random
  phases in x and z, Neumann-compatible ``cos(pi m_y (y+1)/2)`` in y) plus
  ``1e-3 U[-1,1)`` noise, volume-weighted mean removed, ``b = V f``.
* ``batch_system``: system ``i`` of a batch of independent Poisson problems
  on the unit box: shape uniform from 32^3, 48^3, 64^3, 96^3, 128^3 and
  64x32x128; per axis a boundary pair uniform from Neumann(0),
  Dirichlet(0) and periodic; ``f`` as in ``channel`` (cosine modes
  ``0 < |m| <= 8`` with amplitude ``|m|^(-5/3)``, random phases along the
  first and third axes, ``cos(pi m_y y)`` along the second, plus ``1e-3``
  noise), the mean removed when no face is Dirichlet; ``b = V f``.
* ``spheres``: unit cube, all Neumann, density 1000 outside and 1 inside
  ``nspheres`` seeded spheres (radius U[0.03,0.1]) with the interface smoothed
  over 1.5 cells; face coefficient ``lambda`` = arithmetic face average of
  ``1/rho``; ``f = div(u)`` of a seeded face-velocity field with ``u.n = 0`` on
  the walls, ``b = V f`` (sums to zero exactly up to rounding).

Fields are returned in the solver's reference layout: ghost-padded
``(Nx+2, Ny+2, Nz+2)`` float64 with z contiguous.
"""
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
CASE_UNIFORM, CASE_CHANNEL, CASE_SPHERES, CASE_BATCH = 1, 2, 3, 4

NEUMANN, DIRICHLET, PERIODIC = 0, 1, 2


def _mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def stream_base(case, seed, field):
    key = (np.uint64(case) << np.uint64(40)) ^ (np.uint64(seed) << np.uint64(20)) \
        ^ np.uint64(field)
    return _mix64(key)


def raw64(case, seed, field, n, start=0):
    """Draws start..start+n-1 of stream (case, seed, field) as uint64."""
    base = stream_base(case, seed, field)
    i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix64(base + i * GOLDEN)


def uniform_pm1(case, seed, field, n, start=0):
    z = raw64(case, seed, field, n, start)
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0


def uniform01(case, seed, field, n, start=0):
    z = raw64(case, seed, field, n, start)
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class Problem:
    """A Poisson problem in the solver's input form."""

    def __init__(self, name, vertices, bc, b, lam=None, bc_value=(0.0,) * 6):
        self.name = name
        self.vertices = [np.ascontiguousarray(v, np.float64) for v in vertices]
        self.bc = tuple(int(q) for q in bc)
        self.bc_value = tuple(float(q) for q in bc_value)
        self.b = b
        self.lam = lam
        self.n = tuple(len(v) - 1 for v in self.vertices)

    @property
    def padded(self):
        return tuple(q + 2 for q in self.n)


def _pad(inner):
    out = np.zeros(tuple(s + 2 for s in inner.shape))
    out[1:-1, 1:-1, 1:-1] = inner
    return out


def uniform_neumann(n=64, seed=0):
    """Plain all-Neumann Poisson on the unit cube (64^3 CPU case)."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    verts = [np.arange(q + 1, dtype=np.float64) / q for q in (nx, ny, nz)]
    f = uniform_pm1(CASE_UNIFORM, seed, 0, nx * ny * nz).reshape(nx, ny, nz)
    f -= f.mean()
    vol = (1.0 / nx) * (1.0 / ny) * (1.0 / nz)
    return Problem("uniform_neumann", verts, (NEUMANN,) * 6, _pad(f * vol))


def channel_vertices(n=256, beta=2.0):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    vx = np.arange(nx + 1, dtype=np.float64) * (4.0 * math.pi / nx)
    j = np.arange(ny + 1, dtype=np.float64)
    vy = np.tanh(beta * (2.0 * j / ny - 1.0)) / math.tanh(beta)
    vy[0], vy[-1] = -1.0, 1.0
    vz = np.arange(nz + 1, dtype=np.float64) * (2.0 * math.pi / nz)
    return [vx, vy, vz]


def _modes(mmax=8):
    out = []
    for mx in range(0, mmax + 1):
        for my in range(0, mmax + 1):
            for mz in range(0, mmax + 1):
                m2 = mx * mx + my * my + mz * mz
                if 0 < m2 <= mmax * mmax:
                    out.append((mx, my, mz))
    return out


def channel(n=256, seed=0, beta=2.0):
    """Channel-flow pressure Poisson (256^3, periodic x/z, stretched y)."""
    verts = channel_vertices(n, beta)
    nx, ny, nz = (len(v) - 1 for v in verts)
    xc = 0.5 * (verts[0][:-1] + verts[0][1:])
    yc = 0.5 * (verts[1][:-1] + verts[1][1:])
    zc = 0.5 * (verts[2][:-1] + verts[2][1:])
    Lx, Lz = 4.0 * math.pi, 2.0 * math.pi
    modes = _modes(8)
    ph = uniform01(CASE_CHANNEL, seed, 1, 2 * len(modes)) * (2.0 * math.pi)
    # f = sum_m a_m cos(kx x + phx) cos(pi my (y+1)/2) cos(kz z + phz),
    # summed per my as G_my(x,z) = CX^T diag(a) CZ, then over my.
    G = np.zeros((9, nx, nz))
    amp = np.array([(mx * mx + my * my + mz * mz) ** (-5.0 / 6.0)
                    for (mx, my, mz) in modes])
    for my in range(9):
        sel = [q for q, m in enumerate(modes) if m[1] == my]
        if not sel:
            continue
        CX = np.stack([np.cos(2.0 * math.pi * modes[q][0] * xc / Lx + ph[2 * q])
                       for q in sel])
        CZ = np.stack([np.cos(2.0 * math.pi * modes[q][2] * zc / Lz + ph[2 * q + 1])
                       for q in sel])
        G[my] = CX.T @ (amp[sel][:, None] * CZ)
    CY = np.stack([np.cos(math.pi * my * (yc + 1.0) / 2.0) for my in range(9)])
    f = np.einsum("mxz,my->xyz", G, CY, optimize=True)
    f += 1e-3 * uniform_pm1(CASE_CHANNEL, seed, 2, nx * ny * nz).reshape(nx, ny, nz)
    dx, dy, dz = (np.diff(v) for v in verts)
    vol = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    f -= (vol * f).sum() / vol.sum()
    bc = (PERIODIC, PERIODIC, NEUMANN, NEUMANN, PERIODIC, PERIODIC)
    return Problem("channel", verts, bc, _pad(vol * f))


BATCH_SHAPES = ((32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96),
                (128, 128, 128), (64, 32, 128))
BATCH_BCS = (NEUMANN, DIRICHLET, PERIODIC)


def batch_layout(i, seed=0):
    """(shape, six face kinds) of system i of batch seed."""
    z = raw64(CASE_BATCH, seed, 0, 4, start=4 * i)
    shape = BATCH_SHAPES[int(z[0] % np.uint64(len(BATCH_SHAPES)))]
    bc = []
    for d in range(3):
        k = BATCH_BCS[int(z[1 + d] % np.uint64(3))]
        bc += [k, k]
    return shape, tuple(bc)


def batch_system(i, seed=0, shape=None):
    """System i of a batch of independent unit-box Poisson problems
    (``shape`` overrides the drawn shape, e.g. for small test sizes)."""
    drawn, bc = batch_layout(i, seed)
    nx, ny, nz = shape or drawn
    verts = [np.arange(q + 1, dtype=np.float64) / q for q in (nx, ny, nz)]
    xc, yc, zc = ((np.arange(q, dtype=np.float64) + 0.5) / q
                  for q in (nx, ny, nz))
    modes = _modes(8)
    key = 1 + 4 * i
    ph = uniform01(CASE_BATCH, seed, key, 2 * len(modes)) * (2.0 * math.pi)
    amp = np.array([(mx * mx + my * my + mz * mz) ** (-5.0 / 6.0)
                    for (mx, my, mz) in modes])
    G = np.zeros((9, nx, nz))
    for my in range(9):
        sel = [q for q, m in enumerate(modes) if m[1] == my]
        CX = np.stack([np.cos(2.0 * math.pi * modes[q][0] * xc + ph[2 * q])
                       for q in sel])
        CZ = np.stack([np.cos(2.0 * math.pi * modes[q][2] * zc + ph[2 * q + 1])
                       for q in sel])
        G[my] = CX.T @ (amp[sel][:, None] * CZ)
    CY = np.stack([np.cos(math.pi * my * yc) for my in range(9)])
    f = np.einsum("mxz,my->xyz", G, CY, optimize=True)
    f += 1e-3 * uniform_pm1(CASE_BATCH, seed, key + 1,
                            nx * ny * nz).reshape(nx, ny, nz)
    if DIRICHLET not in bc:
        f -= f.mean()
    vol = (1.0 / nx) * (1.0 / ny) * (1.0 / nz)
    return Problem("batch_system", verts, bc, _pad(vol * f))


def _smooth_heaviside(phi, eps):
    """0 for phi < -eps, 1 for phi > eps, smooth in between."""
    t = np.clip(phi / eps, -1.0, 1.0)
    return 0.5 * (1.0 + t + np.sin(math.pi * t) / math.pi)


def spheres(n=192, seed=0, nspheres=32, rho_out=1000.0, rho_in=1.0):
    """Variable-coefficient pressure equation div(1/rho grad p) = f."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    verts = [np.arange(q + 1, dtype=np.float64) / q for q in (nx, ny, nz)]
    h = 1.0 / max(nx, ny, nz)
    c = uniform01(CASE_SPHERES, seed, 1, 3 * nspheres).reshape(nspheres, 3)
    r = 0.03 + 0.07 * uniform01(CASE_SPHERES, seed, 2, nspheres)
    xc = [0.5 * (v[:-1] + v[1:]) for v in verts]
    # padded cell centres (ghosts mirrored) so that face averages exist
    xp = [np.concatenate(([2 * a[0] - a[1]], a, [2 * a[-1] - a[-2]])) for a in xc]
    phi = np.full(tuple(len(a) for a in xp), np.inf)
    # each sphere's signed distance only inside its bounding box grown by
    # 4h: outside it the distance exceeds the smoothing width 1.5h, where
    # the smoothed Heaviside is saturated, so the result is the same as
    # with the distance everywhere
    for s in range(nspheres):
        sl = []
        for d in range(3):
            lo = np.searchsorted(xp[d], c[s, d] - (r[s] + 4 * h), "left")
            hi = np.searchsorted(xp[d], c[s, d] + (r[s] + 4 * h), "right")
            sl.append(slice(lo, hi))
        X = xp[0][sl[0], None, None]
        Y = xp[1][None, sl[1], None]
        Z = xp[2][None, None, sl[2]]
        d = np.sqrt((X - c[s, 0]) ** 2 + (Y - c[s, 1]) ** 2 + (Z - c[s, 2]) ** 2) - r[s]
        sub = phi[sl[0], sl[1], sl[2]]
        np.minimum(sub, d, out=sub)
    Hin = _smooth_heaviside(-phi, 1.5 * h)  # 1 inside spheres
    rho = rho_out + (rho_in - rho_out) * Hin
    inv = 1.0 / rho
    lam = [np.zeros_like(inv) for _ in range(3)]
    # lower-face coefficient of padded cell p: average of cells p-1 and p
    lam[0][1:, :, :] = 0.5 * (inv[:-1, :, :] + inv[1:, :, :])
    lam[1][:, 1:, :] = 0.5 * (inv[:, :-1, :] + inv[:, 1:, :])
    lam[2][:, :, 1:] = 0.5 * (inv[:, :, :-1] + inv[:, :, 1:])
    # divergence of a random face velocity with zero normal flux on walls
    ux = uniform_pm1(CASE_SPHERES, seed, 3, (nx + 1) * ny * nz).reshape(nx + 1, ny, nz)
    uy = uniform_pm1(CASE_SPHERES, seed, 4, nx * (ny + 1) * nz).reshape(nx, ny + 1, nz)
    uz = uniform_pm1(CASE_SPHERES, seed, 5, nx * ny * (nz + 1)).reshape(nx, ny, nz + 1)
    ux[0], ux[-1] = 0.0, 0.0
    uy[:, 0], uy[:, -1] = 0.0, 0.0
    uz[:, :, 0], uz[:, :, -1] = 0.0, 0.0
    dx, dy, dz = (np.diff(v) for v in verts)
    flux = ((ux[1:] - ux[:-1]) * (dy[None, :, None] * dz[None, None, :])
            + (uy[:, 1:] - uy[:, :-1]) * (dx[:, None, None] * dz[None, None, :])
            + (uz[:, :, 1:] - uz[:, :, :-1]) * (dx[:, None, None] * dy[None, :, None]))
    # b = V f = integral of div(u) = net outward flux
    b = flux - flux.mean()
    return Problem("spheres", verts, (NEUMANN,) * 6, _pad(b), lam=lam)
