"""CPU restatement of splitPoissonSolver's modified source, for one process.

Test infrastructure only: the GPU split source (k_split_phi, k_split_div in
briscola_amd/csrc/bmg_kernels.hip) is checked against it, never built on it.

splitPoissonSolver.C:124-137 with the linear face gradient
(linearFaceGradientScheme.C:44-53) and the explicit divergence of a face
flux (divergenceScheme.C:66-95):

* phi_d = (((lambda_d/lambda0) - 1) * g) * A_d on the lower faces normal to
  d, faces 1..N_d+1 along d, interior cells along the others, with
  g = (x[c - e_d] - x[c]) * delta_d, delta_d = 1/|cc_c - cc_{c-e_d}| and A_d
  the product of the two in-face widths;
* div(phi) per interior cell: zero, plus the lower face, minus the upper
  face, directions x, y, z in turn, times 1/((dx*dy)*dz);
* bHat = b/lambda0 + div(phi).

Element-wise numpy operations round exactly as the C expressions they
restate, so the GPU result is compared bit for bit.
"""
import numpy as np


def phi(x, lam, lam0, widths, deltas):
    """The three face arrays (padded layout, zero off the face range)."""
    n = [len(w) - 2 for w in widths]
    out = []
    for d in range(3):
        f = np.zeros_like(x)
        hi = [q + 1 for q in n]
        hi[d] += 1
        sl = tuple(slice(1, h) for h in hi)
        lo = tuple(slice(1 - (e == d), h - (e == d)) for e, h in enumerate(hi))
        idx = [np.arange(1, h) for h in hi]
        shape = lambda e: [-1 if q == e else 1 for q in range(3)]
        g = (x[lo] - x[sl]) * deltas[d][idx[d]].reshape(shape(d))
        a, b = [e for e in range(3) if e != d]
        fa = (widths[a][idx[a]].reshape(shape(a))
              * widths[b][idx[b]].reshape(shape(b)))
        r = lam[d][sl] / lam0 - 1.0
        f[sl] = (r * g) * fa
        out.append(f)
    return out


def source(x, lam, lam0, b, widths, deltas):
    """(bHat, phi): bHat on interior cells (zero ghosts)."""
    ph = phi(x, lam, lam0, widths, deltas)
    n = [len(w) - 2 for w in widths]
    i = [slice(1, q + 1) for q in n]
    acc = np.zeros(tuple(n))
    for d in range(3):
        up = list(i)
        up[d] = slice(2, n[d] + 2)
        acc = acc + ph[d][tuple(i)]
        acc = acc - ph[d][tuple(up)]
    cv = (widths[0][1:-1, None, None] * widths[1][None, 1:-1, None]) \
        * widths[2][None, None, 1:-1]
    icv = 1.0 / cv
    base = b[tuple(i)] / lam0 if b is not None else 0.0
    out = np.zeros_like(x)
    out[tuple(i)] = base + acc * icv
    return out, ph
