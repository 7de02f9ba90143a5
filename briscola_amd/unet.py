"""ctypes binding of the learned preconditioner (``include/bmg_unet.h``) and
its data preparation.

The network is a 4-level 3-D U-Net (16/32/64/128 channels, 3x3x3
convolutions, BF16 weights and activations, FP32 accumulation) whose
forward, data-gradient and weight-gradient kernels run on the gfx950 matrix
cores (``csrc/unet_kernels.hip``).  Buffers passed to the entry points are
device pointers (torch tensors on ``cuda:0`` in the helpers below).

Input scaling (the same in training and in the solver): ``f = b / rms(b)``,
``c = (log10(lambda_mean) + 1.5) / 1.5`` with ``lambda_mean`` the mean of the
six face coefficients of a cell; target ``t = x * h * s / rms(b)`` with ``h``
the level-0 cell width along the first axis and ``s = 10**mean(log10
lambda_mean)`` the geometric-mean coefficient.
"""
import ctypes

import numpy as np

from . import bmg

EXPORTS = [
    "bmg_unet_last_error", "bmg_unet_create", "bmg_unet_destroy",
    "bmg_unet_num_params", "bmg_unet_get_params", "bmg_unet_set_params",
    "bmg_unet_get_grads", "bmg_unet_forward", "bmg_unet_train_step",
    "bmg_unet_op_conv", "bmg_unet_op_conv_dgrad", "bmg_unet_op_conv_wgrad",
    "bmg_set_learned",
]

_sig_done = False


def _lib():
    global _sig_done
    lib = bmg.load()
    if _sig_done:
        return lib
    P, I, F, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_long
    FP = ctypes.POINTER(ctypes.c_float)
    sig = {
        "bmg_unet_last_error": (ctypes.c_char_p, []),
        "bmg_unet_create": (I, [ctypes.POINTER(P), ctypes.c_ulonglong]),
        "bmg_unet_destroy": (I, [P]),
        "bmg_unet_num_params": (L, [P]),
        "bmg_unet_get_params": (I, [P, FP]),
        "bmg_unet_set_params": (I, [P, FP]),
        "bmg_unet_get_grads": (I, [P, FP]),
        "bmg_unet_forward": (I, [P, I, I, I, P, P, P]),
        "bmg_unet_train_step": (I, [P, I, I, I, P, P, P, F, FP]),
        "bmg_unet_op_conv": (I, [I, I, I, P, I, I, P, I, I, I, P, I, P, I]),
        "bmg_unet_op_conv_dgrad": (I, [I, I, I, P, I, I, P, I, I, P, I]),
        "bmg_unet_op_conv_wgrad": (I, [I, I, I, P, I, I, P, I, I, I, I, P]),
        "bmg_set_learned": (I, [P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _sig_done = True
    return lib


def _check(code):
    if code != 0:
        raise bmg.BmgError(code, _lib().bmg_unet_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def round32(k):
    return (k + 31) // 32 * 32


class Net:
    """A U-Net on the device (weights initialised from ``seed``)."""

    def __init__(self, seed=0):
        self.lib = _lib()
        p = ctypes.c_void_p()
        _check(self.lib.bmg_unet_create(ctypes.byref(p), seed))
        self.ptr = p

    def close(self):
        if self.ptr:
            self.lib.bmg_unet_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return int(self.lib.bmg_unet_num_params(self.ptr))

    def params(self):
        a = np.zeros(self.num_params, np.float32)
        _check(self.lib.bmg_unet_get_params(
            self.ptr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.num_params
        _check(self.lib.bmg_unet_set_params(
            self.ptr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def grads(self):
        a = np.zeros(self.num_params, np.float32)
        _check(self.lib.bmg_unet_get_grads(
            self.ptr, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return a

    def forward(self, f, c):
        """f, c: float32 cuda tensors (D, H, W); returns the output tensor."""
        import torch
        out = torch.empty_like(f)
        D, H, W = f.shape
        _check(self.lib.bmg_unet_forward(self.ptr, D, H, W, _ptr(f), _ptr(c),
                                         _ptr(out)))
        return out

    def train_step(self, f, c, target, lr):
        D, H, W = f.shape
        loss = ctypes.c_float()
        _check(self.lib.bmg_unet_train_step(self.ptr, D, H, W, _ptr(f), _ptr(c),
                                            _ptr(target), float(lr),
                                            ctypes.byref(loss)))
        return float(loss.value)

    def attach(self, ctx):
        """Use this network in a solver context (krylov=KRYLOV_FCG_LEARNED)."""
        _check(self.lib.bmg_set_learned(ctx.ptr, self.ptr))


def op_conv(x, w, bias, cout, relu, taps=27):
    """y = act(conv(x, w) + bias) with the library kernel; x (D,H,W,cin)
    bfloat16 cuda tensor, w (cout, kp) bfloat16, bias float32 or None."""
    import torch
    D, H, W, cin = x.shape
    y = torch.empty((D, H, W, cout), dtype=torch.bfloat16, device=x.device)
    _check(_lib().bmg_unet_op_conv(
        D, H, W, _ptr(x), cin, cin, _ptr(w), w.shape[1], taps, cout,
        _ptr(bias) if bias is not None else None, int(relu), _ptr(y), cout))
    return y


def op_conv_dgrad(dz, w, cin):
    import torch
    D, H, W, cout = dz.shape
    dx = torch.empty((D, H, W, cin), dtype=torch.bfloat16, device=dz.device)
    _check(_lib().bmg_unet_op_conv_dgrad(D, H, W, _ptr(dz), cout, cout, _ptr(w),
                                         w.shape[1], cin, _ptr(dx), cin))
    return dx


def op_conv_wgrad(x, dz, kp, chunk=256):
    import torch
    D, H, W, cin = x.shape
    cout = dz.shape[3]
    dw = torch.zeros((cout, kp), dtype=torch.float32, device=x.device)
    _check(_lib().bmg_unet_op_conv_wgrad(D, H, W, _ptr(x), cin, cin, _ptr(dz),
                                         cout, cout, kp, chunk, _ptr(dw)))
    return dw


# ---------------------------------------------------------------------------
# data preparation (host side, numpy)

def coefficient_channel(n, lam):
    """c = (log10(mean of the six face coefficients) + 1.5) / 1.5 per cell;
    lam: three padded lower-face arrays or None (coefficient-free)."""
    if lam is None:
        return np.ones(n, np.float32)
    lx, ly, lz = lam
    m = (lx[1:-1, 1:-1, 1:-1] + lx[2:, 1:-1, 1:-1] + ly[1:-1, 1:-1, 1:-1]
         + ly[1:-1, 2:, 1:-1] + lz[1:-1, 1:-1, 1:-1] + lz[1:-1, 1:-1, 2:]) / 6.0
    v = np.where(m > 0, np.log10(np.maximum(m, 1e-300)), -6.0)
    return ((v + 1.5) / 1.5).astype(np.float32)


def coefficient_scale(c):
    """Geometric mean of the cell coefficients from the channel c."""
    return float(10.0 ** np.mean(1.5 * c.astype(np.float64) - 1.5))


def scaled_pair(prob, x=None):
    """(f, c, t, rms, h*s) of a problem; t only when the solution x (padded)
    is given."""
    b = prob.b[1:-1, 1:-1, 1:-1]
    rms = float(np.sqrt(np.mean(b * b)))
    v0 = prob.vertices[0]
    h = float(v0[-1] - v0[0]) / (len(v0) - 1)
    f = (b / rms).astype(np.float32)
    c = coefficient_channel(prob.n, prob.lam)
    hs = h * coefficient_scale(c)
    t = None
    if x is not None:
        xi = x[1:-1, 1:-1, 1:-1]
        t = ((xi - xi.mean()) * hs / rms).astype(np.float32)
    return f, c, t, rms, hs
