"""Periodic red-black Gauss-Seidel over a whole interior, restated in numpy
for the tests of the wrapped smoothing phases (bmg_set_periodic_smoothing):
ghosts of the periodic axes made current, then per phase the Gauss-Seidel
update of one colour in the order and rounding of the library's kernels,
and the ghosts of the updated cells rewritten.  Test infrastructure only."""
import numpy as np


def fill_periodic(x, axes):
    """Face ghosts of the periodic axes from the cells they mirror (interior
    range of the two other axes; edges and corners untouched)."""
    n = [s - 2 for s in x.shape]
    for d in axes:
        lo = [slice(1, -1)] * 3
        src = [slice(1, -1)] * 3
        lo[d], src[d] = 0, n[d]
        x[tuple(lo)] = x[tuple(src)]
        lo[d], src[d] = n[d] + 1, 1
        x[tuple(lo)] = x[tuple(src)]


def wrapped_rbgs(A, x, b, axes, sweeps):
    """Periodic red-black Gauss-Seidel over the whole interior: ghosts made
    current, then per phase x_c = (b - (A1 x_-i + A2 x_+i + A3 x_-j
    + A4 x_+j + A5 x_-k + A6 x_+k)) / A0 on the cells of colour c, summed in
    that order, and the ghosts of the updated cells rewritten.  Phase colours
    as the library orders them for a region [1, N]: the colours of the
    first, second, second-to-last and last counter positions, a phase equal
    to the one before skipped."""
    x = x.copy()
    n = [s - 2 for s in x.shape]
    fill_periodic(x, axes)
    ii, jj, kk = np.meshgrid(*[np.arange(1, q + 1) for q in n], indexing="ij")
    parity = (ii + jj + kk) & 1
    ss, es = 3, sum(n) + 3
    cols = []
    for c in (ss & 1, (ss + 1) & 1, (es + 1) & 1, es & 1):
        if not cols or cols[-1] != c:
            cols.append(c)
    inner = (slice(1, -1),) * 3
    a = [A[inner + (k,)] for k in range(7)]
    bi = b[inner]
    prev = -1
    for _ in range(sweeps):
        for c in cols:
            if c == prev:
                continue
            prev = c
            p = (a[1] * x[:-2, 1:-1, 1:-1] + a[2] * x[2:, 1:-1, 1:-1]
                 + a[3] * x[1:-1, :-2, 1:-1] + a[4] * x[1:-1, 2:, 1:-1]
                 + a[5] * x[1:-1, 1:-1, :-2] + a[6] * x[1:-1, 1:-1, 2:])
            new = (bi - p) / a[0]
            xi = x[inner]
            m = parity == c
            xi[m] = new[m]
            fill_periodic(x, axes)
    return x
