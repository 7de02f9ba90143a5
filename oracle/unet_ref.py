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


# ---------------------------------------------------------------------------
# The whole network (briscola_amd/csrc/unet.hip), float64, for the gradient
# check of a training step.  Parameter layout of bmg_unet_get_params: for
# each convolution E0, E1, E2, E3, D2, D1, D0 the weights (cout, kp) then
# the bias (cout); then the head's 16 weights and its bias.  Channel order
# of the decoder inputs: [upsampled | skip].
#
# bf16=True rounds to BF16 wherever the device stores BF16 (the packed
# input, every convolution, pooling and upsampling output, the head's input
# gradient, the data gradients, the pooled and upsampled gradients; the
# convolution weights are the BF16 copies of the parameters), so that the
# device result differs from it by FP32 accumulation order only;
# bf16=False is the exact float64 function of the parameters, which the
# tests pin by central differences.

# (cin padded, cout, cin real, level)
CONVS = [(8, 16, 2, 0), (16, 32, 16, 1), (32, 64, 32, 2), (64, 128, 64, 3),
         (192, 64, 192, 2), (96, 32, 96, 1), (48, 16, 48, 0)]


def round32(k):
    return (k + 31) // 32 * 32


def param_layout():
    """[(weight offset, bias offset, kp)] per convolution, head offset,
    total count."""
    out, off = [], 0
    for cin, cout, _, _ in CONVS:
        kp = round32(27 * cin)
        out.append((off, off + cout * kp, kp))
        off += cout * kp + cout
    return out, off, off + 17


def _split(P):
    lay, hoff, n = param_layout()
    assert P.size == n, (P.size, n)
    W, B = [], []
    for (wo, bo, kp), (cin, cout, _, _) in zip(lay, CONVS):
        W.append(np.asarray(P[wo:wo + cout * kp], np.float64).reshape(cout, kp))
        B.append(np.asarray(P[bo:bo + cout], np.float64))
    return W, B, np.asarray(P[hoff:hoff + 17], np.float64)


def _r(a, bf16):
    return to_bf16(a.astype(np.float32)).astype(np.float64) if bf16 else a


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5)) \
        * 0.125


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def up_bwd(dy):
    D, H, W, C = dy.shape
    return dy.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def net_forward(P, f, c, bf16=False):
    """Output (D, H, W) and the activations the backward pass needs."""
    W, B, head = _split(P)
    if bf16:
        W = [_r(w, True) for w in W]
    D, H, Wd = f.shape
    xin = np.zeros((D, H, Wd, 8))
    xin[..., 0] = f
    xin[..., 1] = c
    xin = _r(xin, bf16)
    a = {"xin": xin}
    cv = lambda i, x: _r(conv(x, W[i], B[i], relu=True), bf16)
    e0 = cv(0, xin)
    p1 = _r(pool(e0), bf16)
    e1 = cv(1, p1)
    p2 = _r(pool(e1), bf16)
    e2 = cv(2, p2)
    p3 = _r(pool(e2), bf16)
    e3 = cv(3, p3)
    cat2 = np.concatenate([up(e3), e2], axis=-1)
    d2 = cv(4, cat2)
    cat1 = np.concatenate([up(d2), e1], axis=-1)
    d1 = cv(5, cat1)
    cat0 = np.concatenate([up(d1), e0], axis=-1)
    d0 = cv(6, cat0)
    out = head[16] + d0 @ head[:16]
    a.update(e0=e0, p1=p1, e1=e1, p2=p2, e2=e2, p3=p3, e3=e3, cat2=cat2,
             d2=d2, cat1=cat1, d1=d1, cat0=cat0, d0=d0, W=W)
    return out, a


def net_loss_grad(P, f, c, target, bf16=False):
    """Mean-removed squared error L = mean((o - t - mean(o - t))^2) of the
    network output o and its gradient in the parameter layout."""
    out, a = net_forward(P, f, c, bf16)
    M = out.size
    d = out - target
    d = d - d.mean()
    loss = float(np.mean(d * d))
    g = 2.0 * d / M                                   # dL/d out
    W, B, head = _split(P)
    Wc = a["W"]
    lay, hoff, n = param_layout()
    G = np.zeros(n)
    # head
    G[hoff:hoff + 16] = np.tensordot(g, a["d0"], axes=3)
    G[hoff + 16] = g.sum()
    dd0 = _r(g[..., None] * head[None, None, None, :16], bf16)

    def conv_back(i, x, dz):
        """Weight and bias gradients of convolution i; input gradient."""
        cin, cout, _, _ = CONVS[i]
        wo, bo, kp = lay[i]
        G[wo:wo + cout * kp] = conv_wgrad(x, dz, kp).ravel()
        G[bo:bo + cout] = dz.reshape(-1, cout).sum(axis=0)
        return _r(conv_dgrad(dz, Wc[i], cin), bf16)

    relu = lambda dy, y: np.where(y > 0, dy, 0.0)
    dcat0 = conv_back(6, a["cat0"], relu(dd0, a["d0"]))
    dd1 = _r(up_bwd(dcat0[..., :32]), bf16)
    dcat1 = conv_back(5, a["cat1"], relu(dd1, a["d1"]))
    dd2 = _r(up_bwd(dcat1[..., :64]), bf16)
    dcat2 = conv_back(4, a["cat2"], relu(dd2, a["d2"]))
    de3 = _r(up_bwd(dcat2[..., :128]), bf16)
    dp3 = conv_back(3, a["p3"], relu(de3, a["e3"]))
    de2 = _r(dcat2[..., 128:] + 0.125 * up(dp3), bf16)
    dp2 = conv_back(2, a["p2"], relu(de2, a["e2"]))
    de1 = _r(dcat1[..., 64:] + 0.125 * up(dp2), bf16)
    dp1 = conv_back(1, a["p1"], relu(de1, a["e1"]))
    de0 = _r(dcat0[..., 32:] + 0.125 * up(dp1), bf16)
    conv_back(0, a["xin"], relu(de0, a["e0"]))
    return loss, G
