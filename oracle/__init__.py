"""CPU oracle for the Briscola MG solve -- test infrastructure only.

``Oracle`` wraps ``liboracle.so`` (the C++ restatement in ``bmg_oracle.cpp``)
through ctypes.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
(``briscola_amd``) never does.

Fields use the reference layout: ghost-padded ``(Nx+2, Ny+2, Nz+2)`` float64
arrays, z contiguous (src/briscolaCore/include/utils.H:559-570); stencils are
``(Nx+2, Ny+2, Nz+2, 7)`` arrays of ``[C,L,R,B,T,A,F]``
(src/briscolaCore/primitives/StencilSpace/stencil/stencil.H:21).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

NEUMANN, DIRICHLET, PERIODIC = 0, 1, 2
VCYCLE, WCYCLE, FCYCLE = 0, 1, 2
COARSE_SMOOTH, COARSE_DIRECT = 0, 1
RBGS, JAC, LEXGS = 0, 1, 2
X, B, R = 0, 1, 2

_lib = None


def build(quiet=True):
    """Compile liboracle.so with g++ (-ffp-contract=off keeps the reference's
    rounding sequence)."""
    import subprocess
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC",
           "-fopenmp", "-shared", "-std=c++17", "-o", LIB_PATH,
           os.path.join(_HERE, "bmg_oracle.cpp")]
    subprocess.run(cmd, check=True, capture_output=quiet)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    sig = {
        "oracle_create": (P, [I, D, D, D, I, D]),
        "oracle_destroy": (None, [P]),
        "oracle_set_threads": (None, [c_int]),
        "oracle_get_threads": (c_int, []),
        "oracle_set_opts": (None, [P, c_double, c_double, c_int, c_int, c_int,
                                   c_int, c_int, c_int, c_int, c_int]),
        "oracle_num_levels": (c_int, [P]),
        "oracle_level_shape": (None, [P, c_int, I]),
        "oracle_level_ratio": (None, [P, c_int, I]),
        "oracle_level_centres": (None, [P, c_int, c_int, D]),
        "oracle_level_widths": (None, [P, c_int, c_int, D]),
        "oracle_assemble_laplacian": (None, [P, D, D, D]),
        "oracle_assemble_laplacian_ddt": (None, [P, D, D, D, ctypes.c_double,
                                                 ctypes.c_double]),
        "oracle_set_stencil": (None, [P, c_int, D]),
        "oracle_get_stencil": (None, [P, c_int, D]),
        "oracle_set_symmetric": (None, [P, c_int]),
        "oracle_singular": (c_int, [P]),
        "oracle_set_field": (None, [P, c_int, c_int, D]),
        "oracle_get_field": (None, [P, c_int, c_int, D]),
        "oracle_pin_and_eliminate": (None, [P]),
        "oracle_prepare_coarse": (None, [P]),
        "oracle_smooth": (None, [P, c_int, c_int]),
        "oracle_residual": (None, [P, c_int, c_int]),
        "oracle_restrict": (None, [P, c_int]),
        "oracle_prolong": (None, [P, c_int]),
        "oracle_correct_bcs": (None, [P, c_int]),
        "oracle_coarse_solve": (None, [P]),
        "oracle_cycle": (None, [P]),
        "oracle_correct_eliminated": (None, [P]),
        "oracle_sum_sq_residual": (c_double, [P]),
        "oracle_norm_factor": (None, [P, D]),
        "oracle_solve": (None, [P, c_int, D]),
        "oracle_coarse_dense": (c_int, [P, D, D, I]),
        "oracle_solve_diagonal": (None, [P]),
        "oracle_cycle_counts": (None, [P, c_int, ctypes.POINTER(ctypes.c_long)]),
        "oracle_set_boundary_values": (c_int, [P, c_int, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def set_threads(n):
    """Threads of the oracle's per-cell loops (OpenMP over planes; every
    result is the same for any count)."""
    _load().oracle_set_threads(int(n))


def get_threads():
    return int(_load().oracle_get_threads())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Oracle:
    """One MG problem: mesh hierarchy, operator, fields, solver settings."""

    def __init__(self, vertices, bc, bc_value=(0.0,) * 6):
        lib = _load()
        self.lib = lib
        self.vertices = [np.ascontiguousarray(v, dtype=np.float64)
                         for v in vertices]
        n = np.array([len(v) - 1 for v in self.vertices], dtype=np.int32)
        self.bc = np.array(bc, dtype=np.int32)
        self.bc_value = np.array(bc_value, dtype=np.float64)
        self.ptr = lib.oracle_create(_ip(n), *[_dp(v) for v in self.vertices],
                                     _ip(self.bc), _dp(self.bc_value))
        self.nlevels = lib.oracle_num_levels(self.ptr)
        self.shapes = []
        self.ratios = []
        for l in range(self.nlevels):
            s = np.zeros(3, np.int32)
            lib.oracle_level_shape(self.ptr, l, _ip(s))
            self.shapes.append(tuple(int(q) for q in s))
            lib.oracle_level_ratio(self.ptr, l, _ip(s))
            self.ratios.append(tuple(int(q) for q in s))
        self.set_opts()

    def __del__(self):
        if getattr(self, "ptr", None):
            self.lib.oracle_destroy(self.ptr)
            self.ptr = None

    def padded(self, l):
        return tuple(q + 2 for q in self.shapes[l])

    def set_opts(self, rel_tol=0.0, tolerance=0.0, min_iter=0, max_iter=999,
                 n_sweeps_pre=0, n_sweeps_post=2, cycle=FCYCLE,
                 coarse_mode=COARSE_DIRECT, coarse_level=0, smoother=RBGS):
        self.lib.oracle_set_opts(self.ptr, rel_tol, tolerance, min_iter,
                                 max_iter, n_sweeps_pre, n_sweeps_post, cycle,
                                 coarse_mode, coarse_level, smoother)

    def centres(self, l, axis):
        out = np.zeros(self.shapes[l][axis] + 2)
        self.lib.oracle_level_centres(self.ptr, l, axis, _dp(out))
        return out

    def widths(self, l, axis):
        out = np.zeros(self.shapes[l][axis] + 2)
        self.lib.oracle_level_widths(self.ptr, l, axis, _dp(out))
        return out

    def assemble(self, lam=None, ddt=None):
        """laplacian(lam, x); with ddt = (delta_t, dt_frac) minus the
        Euler time term ddt(x)/dt_frac."""
        dt, frac = ddt if ddt is not None else (0.0, 1.0)
        if lam is None:
            null = ctypes.POINTER(ctypes.c_double)()
            self.lib.oracle_assemble_laplacian_ddt(self.ptr, null, null, null,
                                                   dt, frac)
        else:
            self._lam = [np.ascontiguousarray(a, dtype=np.float64) for a in lam]
            for a in self._lam:
                assert a.shape == self.padded(0)
            self.lib.oracle_assemble_laplacian_ddt(
                self.ptr, *[_dp(a) for a in self._lam], dt, frac)

    def set_stencil(self, l, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        assert A.shape == self.padded(l) + (7,)
        self.lib.oracle_set_stencil(self.ptr, l, _dp(A))

    def stencil(self, l):
        A = np.zeros(self.padded(l) + (7,))
        self.lib.oracle_get_stencil(self.ptr, l, _dp(A))
        return A

    def set_field(self, which, l, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == self.padded(l)
        self.lib.oracle_set_field(self.ptr, which, l, _dp(v))

    def field(self, which, l):
        v = np.zeros(self.padded(l))
        self.lib.oracle_get_field(self.ptr, which, l, _dp(v))
        return v

    def singular(self):
        return bool(self.lib.oracle_singular(self.ptr))

    def pin_and_eliminate(self):
        self.lib.oracle_pin_and_eliminate(self.ptr)

    def prepare_coarse(self):
        self.lib.oracle_prepare_coarse(self.ptr)

    def smooth(self, l, sweeps):
        self.lib.oracle_smooth(self.ptr, l, sweeps)

    def residual(self, l, corr_bcs=True):
        self.lib.oracle_residual(self.ptr, l, int(corr_bcs))

    def restrict(self, l):
        self.lib.oracle_restrict(self.ptr, l)

    def prolong(self, l):
        self.lib.oracle_prolong(self.ptr, l)

    def correct_bcs(self, l):
        self.lib.oracle_correct_bcs(self.ptr, l)

    def coarse_solve(self):
        self.lib.oracle_coarse_solve(self.ptr)

    def cycle(self):
        self.lib.oracle_cycle(self.ptr)

    COUNT_NAMES = ("visits", "smooth", "residual", "restrict", "prolong",
                   "coarse_solve")

    def cycle_counts(self):
        """Per-level operation counts of the last cycle (oracle_cycle or
        the last iteration of solve): dict name -> list over levels."""
        out = {}
        for w, name in enumerate(self.COUNT_NAMES):
            a = np.zeros(self.nlevels, np.int64)
            self.lib.oracle_cycle_counts(
                self.ptr, w, a.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
            out[name] = [int(q) for q in a]
        return out

    def boundary_shape(self, face):
        """Shape of the per-cell boundary data of a face: the level-0
        boundary layer with the normal axis of size 1."""
        s = list(self.shapes[0])
        s[face // 2] = 1
        return tuple(s)

    def set_boundary_values(self, face, values):
        """Per-cell Dirichlet value / Neumann gradient of a face, in the
        reference's patch-block layout (boundary_shape(face), last axis
        fastest); None: the uniform value given at construction."""
        if values is None:
            null = ctypes.POINTER(ctypes.c_double)()
            self.lib.oracle_set_boundary_values(self.ptr, face, null)
            return
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == self.boundary_shape(face), (v.shape, face)
        self._bcarr = getattr(self, "_bcarr", {})
        self._bcarr[face] = v
        rc = self.lib.oracle_set_boundary_values(self.ptr, face, _dp(v))
        assert rc == 0

    def sum_sq_residual(self):
        return self.lib.oracle_sum_sq_residual(self.ptr)

    def norm_factor(self):
        nf = np.zeros(1)
        self.lib.oracle_norm_factor(self.ptr, _dp(nf))
        return float(nf[0])

    def solve(self, const_matrix=False):
        st = np.zeros(3)
        self.lib.oracle_solve(self.ptr, int(const_matrix), _dp(st))
        return int(st[0]), float(st[1]), float(st[2])

    def solve_diagonal(self):
        self.lib.oracle_solve_diagonal(self.ptr)

    def coarse_dense(self):
        l = self.nlevels - 1
        n = int(np.prod(self.shapes[l])) if l >= 0 else 0
        n = max(n, max(int(np.prod(s)) for s in self.shapes))
        dense = np.zeros(n * n)
        lu = np.zeros(n * n)
        piv = np.zeros(n, np.int32)
        m = self.lib.oracle_coarse_dense(self.ptr, _dp(dense), _dp(lu), _ip(piv))
        return (dense[:m * m].reshape(m, m), lu[:m * m].reshape(m, m),
                piv[:m].copy())
