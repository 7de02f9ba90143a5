"""ctypes binding of ``libbmg.so`` (the C ABI declared in ``include/bmg.h``).

There is no fallback: if the HIP library is missing or no GPU is visible the
calls raise.  Arrays use the reference layout -- ghost-padded
``(Nx+2, Ny+2, Nz+2)`` float64 blocks, z contiguous; stencils carry a trailing
axis of 7 ``[C,L,R,B,T,A,F]`` coefficients.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmg.so")

NEUMANN, DIRICHLET, PERIODIC = 0, 1, 2
VCYCLE, WCYCLE, FCYCLE = 0, 1, 2
COARSE_SMOOTH, COARSE_DIRECT = 0, 1
RBGS, JAC, LEXGS = 0, 1, 2
FP64, MIXED = 0, 1
KRYLOV_NONE, KRYLOV_FCG, KRYLOV_FCG_LEARNED = 0, 1, 2
X, B, R = 0, 1, 2
TIME_PHASE, TIME_RESIDUAL, TIME_RESTRICT, TIME_PROLONG, TIME_ITERATION = range(5)

EXPORTS = [
    "bmg_default_opts", "bmg_last_error", "bmg_device_count", "bmg_create",
    "bmg_destroy", "bmg_set_opts", "bmg_num_levels", "bmg_level_shape",
    "bmg_set_operator_stencil", "bmg_set_operator_stencil_device",
    "bmg_set_operator_laplacian", "bmg_set_operator_laplacian_ddt",
    "bmg_get_operator_stencil", "bmg_solve",
    "bmg_solve_device", "bmg_upload_field", "bmg_download_field",
    "bmg_op_prepare", "bmg_op_smooth", "bmg_op_residual", "bmg_op_restrict",
    "bmg_op_prolong", "bmg_op_correct_bcs", "bmg_op_coarse_solve",
    "bmg_op_cycle", "bmg_op_sum_sq_residual", "bmg_time_smoother",
    "bmg_set_graphs", "bmg_set_fused_smoothing",
    "bmg_set_periodic_smoothing", "bmg_solve_diagonal",
    "bmg_batch_create",
    "bmg_batch_destroy", "bmg_solve_batch", "bmg_batch_info",
    "bmg_batch_set_autotune", "bmg_face_flux", "bmg_face_flux_device",
    "bmg_fft_create", "bmg_fft_destroy", "bmg_fft_solve",
    "bmg_fft_solve_device", "bmg_fft_solve_ddt", "bmg_fft_solve_ddt_device",
    "bmg_split_source", "bmg_split_source_device",
    "bmg_set_boundary_values", "bmg_set_boundary_values_device",
    "bmg_renew_source_terms", "bmg_time_op", "bmg_face_flux_noalloc",
    "bmg_batch_tuning", "bmg_batch_set_tile_shape", "bmg_deflation_vectors",
]


class BmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bmg error {code}: {msg}")
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3),
                ("vertices", ctypes.POINTER(ctypes.c_double) * 3),
                ("bc", ctypes.c_int * 6),
                ("bc_value", ctypes.c_double * 6)]


class Opts(ctypes.Structure):
    _fields_ = [("rel_tol", ctypes.c_double), ("tolerance", ctypes.c_double),
                ("min_iter", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("n_sweeps_pre", ctypes.c_int), ("n_sweeps_post", ctypes.c_int),
                ("cycle", ctypes.c_int), ("coarse_mode", ctypes.c_int),
                ("coarse_level", ctypes.c_int), ("smoother", ctypes.c_int),
                ("precision", ctypes.c_int), ("krylov", ctypes.c_int),
                ("deflation", ctypes.c_int), ("reserved", ctypes.c_int * 3)]


class Stats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("status", ctypes.c_int),
                ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double),
                ("solve_ms", ctypes.c_double),
                ("learned_iters", ctypes.c_int),
                ("learned_fallback", ctypes.c_int),
                ("learned_guess_residual", ctypes.c_double),
                ("reserved", ctypes.c_double * 2)]


_lib = None


def load():
    """Load libbmg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run python briscola_amd/build.py")
    # PyTorch for ROCm bundles its own HIP runtime under the soname this
    # library links; the runtime loaded first serves both, and torch finds
    # no device on the system one.  Where torch is installed it is therefore
    # loaded first, whatever order the caller imports the two in.
    if "torch" not in sys.modules:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int)
    sig = {
        "bmg_default_opts": (None, [ctypes.POINTER(Opts)]),
        "bmg_last_error": (ctypes.c_char_p, []),
        "bmg_device_count": (I, [IP]),
        "bmg_create": (I, [ctypes.POINTER(P), ctypes.POINTER(Grid),
                           ctypes.POINTER(Opts)]),
        "bmg_destroy": (I, [P]),
        "bmg_set_opts": (I, [P, ctypes.POINTER(Opts)]),
        "bmg_num_levels": (I, [P]),
        "bmg_level_shape": (I, [P, I, IP]),
        "bmg_set_operator_stencil": (I, [P, I, DP, I]),
        "bmg_set_operator_stencil_device": (I, [P, I, P, I]),
        "bmg_set_operator_laplacian": (I, [P, DP, DP, DP]),
        "bmg_set_operator_laplacian_ddt": (I, [P, DP, DP, DP, ctypes.c_double,
                                               ctypes.c_double]),
        "bmg_get_operator_stencil": (I, [P, I, DP]),
        "bmg_solve": (I, [P, DP, DP, I, ctypes.POINTER(Stats)]),
        "bmg_solve_device": (I, [P, P, P, I, ctypes.POINTER(Stats)]),
        "bmg_upload_field": (I, [P, I, I, DP]),
        "bmg_download_field": (I, [P, I, I, DP]),
        "bmg_op_prepare": (I, [P]),
        "bmg_op_smooth": (I, [P, I, I]),
        "bmg_op_residual": (I, [P, I, I]),
        "bmg_op_restrict": (I, [P, I]),
        "bmg_op_prolong": (I, [P, I]),
        "bmg_op_correct_bcs": (I, [P, I]),
        "bmg_op_coarse_solve": (I, [P]),
        "bmg_op_cycle": (I, [P]),
        "bmg_op_sum_sq_residual": (I, [P, DP]),
        "bmg_time_smoother": (I, [P, I, I, I, DP, DP]),
        "bmg_time_op": (I, [P, I, I, I, I, DP, DP]),
        "bmg_set_graphs": (I, [P, I]),
        "bmg_set_fused_smoothing": (I, [P, I]),
        "bmg_set_periodic_smoothing": (I, [P, I]),
        "bmg_solve_diagonal": (I, [P, DP, DP, DP]),
        "bmg_batch_create": (I, [ctypes.POINTER(P), ctypes.POINTER(P), I]),
        "bmg_batch_destroy": (I, [P]),
        "bmg_solve_batch": (I, [P, ctypes.POINTER(P), ctypes.POINTER(P), I,
                                ctypes.POINTER(Stats)]),
        "bmg_batch_info": (I, [P, IP, ctypes.POINTER(ctypes.c_long)]),
        "bmg_batch_set_autotune": (I, [P, I]),
        "bmg_batch_tuning": (I, [P, IP, IP]),
        "bmg_batch_set_tile_shape": (I, [P, I]),
        "bmg_deflation_vectors": (I, [P, IP]),
        "bmg_face_flux": (I, [P, DP, DP, DP, DP, DP, DP, DP]),
        "bmg_face_flux_noalloc": (I, [P, DP, DP, DP, DP, DP, DP, DP]),
        "bmg_face_flux_device": (I, [P, P, P, P, P, P, P, P]),
        "bmg_fft_create": (I, [ctypes.POINTER(P), ctypes.POINTER(Grid)]),
        "bmg_fft_destroy": (I, [P]),
        "bmg_fft_solve": (I, [P, DP, DP]),
        "bmg_fft_solve_device": (I, [P, P, P]),
        "bmg_fft_solve_ddt": (I, [P, DP, DP, DP, D, D]),
        "bmg_split_source": (I, [P, DP, DP, DP, DP, D, DP, DP, DP, DP, DP]),
        "bmg_split_source_device": (I, [P, P, P, P, P, D, P, P, P, P, P]),
        "bmg_fft_solve_ddt_device": (I, [P, P, P, P, D, D]),
        "bmg_set_boundary_values": (I, [P, I, DP]),
        "bmg_set_boundary_values_device": (I, [P, I, P]),
        "bmg_renew_source_terms": (I, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(code):
    if code != 0:
        raise BmgError(code, load().bmg_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_opts(**kw):
    o = Opts()
    load().bmg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Context:
    """One solver instance: hierarchy, operator, device fields."""

    def __init__(self, vertices, bc, bc_value=(0.0,) * 6, **opts):
        lib = load()
        self.lib = lib
        self._verts = [np.ascontiguousarray(v, dtype=np.float64) for v in vertices]
        g = Grid()
        for d in range(3):
            g.n[d] = len(self._verts[d]) - 1
            g.vertices[d] = _dp(self._verts[d])
        for f in range(6):
            g.bc[f] = int(bc[f])
            g.bc_value[f] = float(bc_value[f])
        self.opts = default_opts(**opts)
        ptr = ctypes.c_void_p()
        _check(lib.bmg_create(ctypes.byref(ptr), ctypes.byref(g),
                              ctypes.byref(self.opts)))
        self.ptr = ptr
        self.nlevels = lib.bmg_num_levels(ptr)
        self.shapes = []
        for l in range(self.nlevels):
            s = (ctypes.c_int * 3)()
            _check(lib.bmg_level_shape(ptr, l, s))
            self.shapes.append(tuple(s))

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.bmg_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def padded(self, l):
        return tuple(q + 2 for q in self.shapes[l])

    def set_opts(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.opts, k):
                raise KeyError(k)
            setattr(self.opts, k, v)
        _check(self.lib.bmg_set_opts(self.ptr, ctypes.byref(self.opts)))

    def boundary_shape(self, face):
        """Shape of a face's per-cell boundary data: the level-0 boundary
        layer with the normal axis of size 1."""
        s = list(self.shapes[0])
        s[face // 2] = 1
        return tuple(s)

    def set_boundary_values(self, face, values):
        """Per-cell Dirichlet value / Neumann gradient of a face (the
        reference's patch block, boundary_shape(face), last axis fastest);
        None: the uniform value of the constructor."""
        if values is None:
            null = ctypes.POINTER(ctypes.c_double)()
            _check(self.lib.bmg_set_boundary_values(self.ptr, face, null))
            return
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != self.boundary_shape(face):
            raise ValueError(f"boundary data shape {v.shape} != "
                             f"{self.boundary_shape(face)}")
        _check(self.lib.bmg_set_boundary_values(self.ptr, face, _dp(v)))

    def renew_source_terms(self):
        """The next solve folds the boundary terms into its b again, as for
        a newly assembled system with the same operator."""
        _check(self.lib.bmg_renew_source_terms(self.ptr))

    def set_stencil(self, l, A, symmetric=True):
        A = np.ascontiguousarray(A, dtype=np.float64)
        if A.shape != self.padded(l) + (7,):
            raise ValueError(f"stencil shape {A.shape} != {self.padded(l) + (7,)}")
        _check(self.lib.bmg_set_operator_stencil(self.ptr, l, _dp(A), int(symmetric)))

    def set_laplacian(self, lam=None, ddt=None):
        """laplacian(lam, x) on every level; ddt = (delta_t, dt_frac) also
        subtracts the Euler time term ddt(x)/dt_frac from the matrix."""
        if lam is None:
            null = ctypes.POINTER(ctypes.c_double)()
            ptrs = [null, null, null]
        else:
            self._lam = [np.ascontiguousarray(a, dtype=np.float64) for a in lam]
            for a in self._lam:
                if a.shape != self.padded(0):
                    raise ValueError("lambda arrays must have the padded level-0 shape")
            ptrs = [_dp(a) for a in self._lam]
        if ddt is None:
            _check(self.lib.bmg_set_operator_laplacian(self.ptr, *ptrs))
        else:
            dt, frac = ddt
            _check(self.lib.bmg_set_operator_laplacian_ddt(
                self.ptr, *ptrs, float(dt), float(frac)))

    def stencil(self, l):
        A = np.zeros(self.padded(l) + (7,))
        _check(self.lib.bmg_get_operator_stencil(self.ptr, l, _dp(A)))
        return A

    def solve(self, b, x=None, const_matrix=False):
        """MG::solve on host arrays; b and x are updated in place."""
        if x is None:
            x = np.zeros(self.padded(0))
        for a in (b, x):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape != self.padded(0):
                raise ValueError("b and x must be C-contiguous float64 padded level-0 arrays")
        st = Stats()
        _check(self.lib.bmg_solve(self.ptr, _dp(b), _dp(x), int(const_matrix),
                                  ctypes.byref(st)))
        return x, st

    def solve_diagonal(self, A0, b, x):
        """MG::solve of a diagonal system (level-0 stencil A0)."""
        A0 = np.ascontiguousarray(A0, dtype=np.float64)
        _check(self.lib.bmg_solve_diagonal(self.ptr, _dp(A0), _dp(b), _dp(x)))
        return x

    def solve_device(self, b_ptr, x_ptr, const_matrix=False):
        """MG::solve on device pointers (e.g. torch tensor .data_ptr())."""
        st = Stats()
        _check(self.lib.bmg_solve_device(self.ptr, ctypes.c_void_p(b_ptr),
                                         ctypes.c_void_p(x_ptr), int(const_matrix),
                                         ctypes.byref(st)))
        return st

    # level operations (kernel-level parity)
    def upload(self, which, l, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.shape != self.padded(l):
            raise ValueError("field shape mismatch")
        _check(self.lib.bmg_upload_field(self.ptr, which, l, _dp(v)))

    def download(self, which, l):
        v = np.zeros(self.padded(l))
        _check(self.lib.bmg_download_field(self.ptr, which, l, _dp(v)))
        return v

    def prepare(self):
        _check(self.lib.bmg_op_prepare(self.ptr))

    def smooth(self, l, sweeps):
        _check(self.lib.bmg_op_smooth(self.ptr, l, sweeps))

    def residual(self, l, corr_bcs=True):
        _check(self.lib.bmg_op_residual(self.ptr, l, int(corr_bcs)))

    def restrict(self, l):
        _check(self.lib.bmg_op_restrict(self.ptr, l))

    def prolong(self, l):
        _check(self.lib.bmg_op_prolong(self.ptr, l))

    def correct_bcs(self, l):
        _check(self.lib.bmg_op_correct_bcs(self.ptr, l))

    def coarse_solve(self):
        _check(self.lib.bmg_op_coarse_solve(self.ptr))

    def cycle(self):
        _check(self.lib.bmg_op_cycle(self.ptr))

    def time_smoother(self, level=0, precision=FP64, reps=50):
        """Average duration (ms) and algorithmic bytes of one red-black phase
        launch on ``level`` (HIP events on the solver's stream)."""
        ms = ctypes.c_double(0.0)
        by = ctypes.c_double(0.0)
        _check(self.lib.bmg_time_smoother(self.ptr, level, precision, reps,
                                          ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def time_op(self, op, level=0, precision=FP64, reps=20):
        """Average duration (ms) and algorithmic bytes of one launch of a
        level operation (TIME_PHASE, TIME_RESIDUAL, TIME_RESTRICT,
        TIME_PROLONG) or of one whole MG iteration (TIME_ITERATION), timed
        with HIP events on the solver's stream."""
        ms = ctypes.c_double(0.0)
        by = ctypes.c_double(0.0)
        _check(self.lib.bmg_time_op(self.ptr, op, level, precision, reps,
                                    ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def split_source(self, x, lam, lambda0, b=None):
        """splitPoissonSolver's correction source on host arrays: returns
        (bHat, [phi_x, phi_y, phi_z])."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in [x] + list(lam)]
        bb = np.ascontiguousarray(b, dtype=np.float64) if b is not None else None
        for a in arrs + ([bb] if bb is not None else []):
            if a.shape != self.padded(0):
                raise ValueError("fields must have the padded level-0 shape")
        bhat = np.zeros(self.padded(0))
        phi = [np.zeros(self.padded(0)) for _ in range(3)]
        _check(self.lib.bmg_split_source(
            self.ptr, *[_dp(a) for a in arrs], float(lambda0),
            _dp(bb) if bb is not None else None, _dp(bhat),
            *[_dp(p) for p in phi]))
        return bhat, phi

    def face_flux(self, x, lam=None):
        """PoissonSolver flux of a level-0 field x (host array, reference
        layout): three padded lower-face arrays."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = [np.zeros(self.padded(0)) for _ in range(3)]
        if lam is None:
            ls = [None, None, None]
        else:
            lam = [np.ascontiguousarray(a, dtype=np.float64) for a in lam]
            ls = [_dp(a) for a in lam]
        _check(self.lib.bmg_face_flux(self.ptr, _dp(x), *ls,
                                      *[_dp(a) for a in out]))
        return out

    def deflation_vectors(self):
        """Number of deflation vectors of the current operator."""
        n = ctypes.c_int(0)
        _check(self.lib.bmg_deflation_vectors(self.ptr, ctypes.byref(n)))
        return n.value

    def set_graphs(self, enable):
        _check(self.lib.bmg_set_graphs(self.ptr, int(enable)))

    def set_fused_smoothing(self, mode):
        """0 (default): one launch per red-black phase; 1: phase pairs in one
        LDS-staged launch on large levels without periodic faces; 2: on every
        level.  Same bits in every mode."""
        _check(self.lib.bmg_set_fused_smoothing(self.ptr, int(mode)))

    def set_periodic_smoothing(self, mode):
        """Red-black smoothing on levels with periodic faces.  1 (default):
        in the FP32 cycles each phase covers the whole interior and writes
        the periodic ghosts with their cells (even cell counts along the
        periodic axes); 2: in the FP64 cycles too (not the reference's
        order); 0: interior phases, ghost fill and boundary slabs (the
        reference's order) everywhere."""
        _check(self.lib.bmg_set_periodic_smoothing(self.ptr, int(mode)))

    def sum_sq_residual(self):
        out = np.zeros(1)
        _check(self.lib.bmg_op_sum_sq_residual(self.ptr, _dp(out)))
        return float(out[0])


class Batch:
    """Independent systems solved together (``bmg_solve_batch``): every
    kernel of an iteration is one launch over all of them.  ``contexts``
    are :class:`Context` objects with their operators set; they must outlive
    the batch."""

    def __init__(self, contexts, autotune=True, tile_shape=0):
        self.lib = load()
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(
            *[c.ptr.value for c in self.contexts])
        p = ctypes.c_void_p()
        _check(self.lib.bmg_batch_create(ctypes.byref(p), arr,
                                         len(self.contexts)))
        self.ptr = p
        if not autotune:
            _check(self.lib.bmg_batch_set_autotune(self.ptr, 0))
        if tile_shape:
            _check(self.lib.bmg_batch_set_tile_shape(self.ptr, tile_shape))

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.bmg_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_device(self, b_ptrs, x_ptrs, const_matrix=False):
        """Device pointers (reference layout) per system; returns the list
        of per-system :class:`Stats`."""
        n = len(self.contexts)
        assert len(b_ptrs) == n and len(x_ptrs) == n
        bp = (ctypes.c_void_p * n)(*b_ptrs)
        xp = (ctypes.c_void_p * n)(*x_ptrs)
        st = (Stats * n)()
        _check(self.lib.bmg_solve_batch(self.ptr, bp, xp, int(const_matrix),
                                        st))
        return list(st)

    def info(self):
        """(launches, virtual workgroups) of one merged iteration."""
        n = ctypes.c_int(0)
        w = ctypes.c_long(0)
        _check(self.lib.bmg_batch_info(self.ptr, ctypes.byref(n),
                                       ctypes.byref(w)))
        return n.value, w.value

    def tuning(self):
        """(launches, launches the autotuner moved to another tile shape)
        of the last solve's merged iteration."""
        n = ctypes.c_int(0)
        k = ctypes.c_int(0)
        _check(self.lib.bmg_batch_tuning(self.ptr, ctypes.byref(n),
                                         ctypes.byref(k)))
        return n.value, k.value


class FFT:
    """The FFT Poisson solver (``bmg_fft_*``): div(grad x) = -b on a grid
    with at least two uniform directions."""

    def __init__(self, vertices, bc, bc_value=(0.0,) * 6):
        self.lib = load()
        self._verts = [np.ascontiguousarray(v, dtype=np.float64) for v in vertices]
        g = Grid()
        for d in range(3):
            g.n[d] = len(self._verts[d]) - 1
            g.vertices[d] = _dp(self._verts[d])
        for f in range(6):
            g.bc[f] = int(bc[f])
            g.bc_value[f] = float(bc_value[f])
        self.shape = tuple(len(v) + 1 for v in self._verts)
        p = ctypes.c_void_p()
        _check(self.lib.bmg_fft_create(ctypes.byref(p), ctypes.byref(g)))
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.bmg_fft_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, x=None, ddt=None, x_old=None):
        """Host arrays (reference layout); x (initial ghosts) is updated in
        place and returned.  ddt = (delta_t, dt_frac) with x_old adds the
        implicit time term."""
        if x is None:
            x = np.zeros(self.shape)
        for a in (b, x, x_old):
            if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous
                                  or a.shape != self.shape):
                raise ValueError("b, x and x_old must be C-contiguous float64 "
                                 "padded arrays")
        bp = _dp(b) if b is not None else None
        if ddt is None:
            _check(self.lib.bmg_fft_solve(self.ptr, bp, _dp(x)))
        else:
            _check(self.lib.bmg_fft_solve_ddt(
                self.ptr, bp, _dp(x), _dp(x_old) if x_old is not None else None,
                float(ddt[0]), float(ddt[1])))
        return x

    def solve_device(self, b_ptr, x_ptr, ddt=None, x_old_ptr=None):
        if ddt is None:
            _check(self.lib.bmg_fft_solve_device(self.ptr, ctypes.c_void_p(b_ptr),
                                                 ctypes.c_void_p(x_ptr)))
        else:
            _check(self.lib.bmg_fft_solve_ddt_device(
                self.ptr, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr),
                ctypes.c_void_p(x_old_ptr), float(ddt[0]), float(ddt[1])))
