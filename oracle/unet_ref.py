"""CPU reference of the learned preconditioner's convolution kernels -- test
infrastructure only (the checker of the device kernels; nothing in the
product imports it).

Layouts follow the kernels: channels-last activations (D, H, W, C); weights
(Cout, Kp) with k = t*Cin + c, t = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding
outside the grid.  BF16 values are held as float32 arrays whose values are
exactly representable in BF16; sums are taken in float64, so a device result
differs from these by its FP32 accumulation and BF16 output rounding only.

There is no reference counterpart (Briscola has no learned component): the
convolution, its data gradient and its weight gradient are pinned by the
definitions below, checked in tests against each other (adjoint identity
<dY, conv(x)> = <dgrad(dY), x> and <dY, conv(x; w)> = <wgrad(x, dY), w>) and
against a finite-difference derivative.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(a):
    """BF16 bit patterns (uint16) of float32 values already rounded."""
    a = np.ascontiguousarray(a, np.float32)
    return (a.view(np.uint32) >> 16).astype(np.uint16)


def taps():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
            for dx in (-1, 0, 1)]


def shifted(x, dz, dy, dx):
    """y[z,y,x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    D, H, W = x.shape[:3]
    y = np.zeros_like(x)
    zs = slice(max(0, -dz), D - max(0, dz))
    ys = slice(max(0, -dy), H - max(0, dy))
    xs = slice(max(0, -dx), W - max(0, dx))
    zt = slice(max(0, dz), D - max(0, -dz))
    yt = slice(max(0, dy), H - max(0, -dy))
    xt = slice(max(0, dx), W - max(0, -dx))
    y[zs, ys, xs] = x[zt, yt, xt]
    return y


def conv(x, w, bias=None, relu=False):
    """3x3x3 convolution: y[m, n] = sum_t sum_c x[m + d_t, c] w[n, t*Cin + c]."""
    D, H, W, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((D, H, W, cout))
    for t, (dz, dy, dx) in enumerate(taps()):
        xs = shifted(x.astype(np.float64), dz, dy, dx)
        wt = w[:, t * cin:(t + 1) * cin].astype(np.float64)      # (cout, cin)
        y += xs @ wt.T
    if bias is not None:
        y += bias[None, None, None, :]
    if relu:
        y = np.maximum(y, 0.0)
    return y


def conv_dgrad(dz, w, cin):
    """Data gradient: dx[m, c] = sum_t sum_n dz[m - d_t, n] w[n, t*Cin + c]."""
    D, H, W, cout = dz.shape
    dx_ = np.zeros((D, H, W, cin))
    for t, (a, b, c) in enumerate(taps()):
        ds = shifted(dz.astype(np.float64), -a, -b, -c)
        wt = w[:, t * cin:(t + 1) * cin].astype(np.float64)      # (cout, cin)
        dx_ += ds @ wt
    return dx_


def conv_wgrad(x, dz, kp):
    """Weight gradient: dw[n, t*Cin + c] = sum_m dz[m, n] x[m + d_t, c]."""
    D, H, W, cin = x.shape
    cout = dz.shape[3]
    dw = np.zeros((cout, kp))
    dzf = dz.reshape(-1, cout).astype(np.float64)
    for t, (a, b, c) in enumerate(taps()):
        xs = shifted(x.astype(np.float64), a, b, c).reshape(-1, cin)
        dw[:, t * cin:(t + 1) * cin] = dzf.T @ xs
    return dw
