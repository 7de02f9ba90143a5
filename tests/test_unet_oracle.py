"""The CPU reference of the learned preconditioner's convolution kernels
(oracle/unet_ref.py) pinned by identities that fix it: the data gradient is
the adjoint of the convolution, the weight gradient is the derivative of
<dY, conv(x; w)> in w, both checked against each other and against central
differences, and BF16 rounding against hand-computed values."""
import numpy as np

from oracle import unet_ref as U


def rng(seed):
    return np.random.default_rng(seed)


def test_bf16_rounding_known_values():
    a = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 0.0],
                 np.float32)
    r = U.to_bf16(a)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1.005859375 =
    # 1 + 1.5*2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125)
    assert abs(r[3] - 3.140625) < 1e-7
    assert U.bf16_bits(np.array([1.0], np.float32))[0] == 0x3F80
    assert r[5] == 0.0


def test_shift_is_zero_padded_translation():
    x = np.arange(4 * 5 * 6, dtype=np.float64).reshape(4, 5, 6, 1)
    y = U.shifted(x, 1, -1, 0)
    assert y[0, 1, 0, 0] == x[1, 0, 0, 0]
    assert y[3].sum() == 0.0 and y[:, 0].sum() == 0.0


def test_conv_identity_kernel():
    """Only the centre tap (t = 13) set to the identity: y = x."""
    x = rng(0).standard_normal((4, 4, 6, 8))
    kp = 27 * 8
    w = np.zeros((8, kp))
    w[:, 13 * 8:14 * 8] = np.eye(8)
    assert np.allclose(U.conv(x, w), x)


def test_dgrad_is_adjoint_of_conv():
    r = rng(1)
    x = r.standard_normal((4, 5, 6, 8))
    dy = r.standard_normal((4, 5, 6, 16))
    w = r.standard_normal((16, 27 * 8))
    lhs = np.sum(dy * U.conv(x, w))
    rhs = np.sum(U.conv_dgrad(dy, w, 8) * x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_wgrad_is_derivative_in_w():
    r = rng(2)
    x = r.standard_normal((4, 4, 4, 8))
    dy = r.standard_normal((4, 4, 4, 16))
    kp = 27 * 8 + 8          # padded K columns stay zero
    w = r.standard_normal((16, kp))
    w[:, 27 * 8:] = 0.0
    g = U.conv_wgrad(x, dy, kp)
    assert np.all(g[:, 27 * 8:] == 0.0)
    # <dy, conv(x; w)> is linear in w: its gradient is g exactly
    assert abs(np.sum(dy * U.conv(x, w[:, :216])) - np.sum(g * w)) < 1e-9 * np.abs(g * w).sum()
    # and a central difference of one coefficient
    i, k = 3, 5 * 8 + 2
    e = 1e-3
    wp, wm = w.copy(), w.copy()
    wp[i, k] += e
    wm[i, k] -= e
    fd = (np.sum(dy * U.conv(x, wp[:, :216])) -
          np.sum(dy * U.conv(x, wm[:, :216]))) / (2 * e)
    assert abs(fd - g[i, k]) < 1e-6 * max(1.0, abs(g[i, k]))


def _net_params(seed):
    """Parameters in the device layout: He-normal weights on the real
    input channels, small random biases, random head."""
    r = rng(seed)
    lay, hoff, n = U.param_layout()
    P = np.zeros(n)
    for (wo, bo, kp), (cin, cout, creal, _) in zip(lay, U.CONVS):
        w = np.zeros((cout, kp))
        for t in range(27):
            w[:, t * cin:t * cin + creal] = (r.standard_normal((cout, creal))
                                             * np.sqrt(2.0 / (27 * creal)))
        P[wo:wo + cout * kp] = w.ravel()
        P[bo:bo + cout] = 0.05 * r.standard_normal(cout)
    P[hoff:hoff + 16] = 0.25 * r.standard_normal(16)
    P[hoff + 16] = 0.1
    return P


def test_network_gradient_central_differences():
    """The float64 U-Net loss gradient (oracle/unet_ref.py) against central
    differences of the loss along random directions confined to single
    layers (first encoder weights, deepest encoder weights, a decoder bias,
    the head) and along one over all parameters."""
    P = _net_params(0)
    r = rng(1)
    D = 8
    f = r.standard_normal((D, D, D))
    c = r.uniform(-1, 1, (D, D, D))
    t = r.standard_normal((D, D, D))
    L, G = U.net_loss_grad(P, f, c, t)
    lay, hoff, n = U.param_layout()

    def loss(Q):
        o, _ = U.net_forward(Q, f, c)
        d = o - t
        d = d - d.mean()
        return float(np.mean(d * d))

    assert abs(loss(P) - L) <= 1e-14 * L
    groups = [slice(lay[0][0], lay[0][1]), slice(lay[3][0], lay[3][1]),
              slice(lay[4][1], lay[4][1] + 64), slice(hoff, hoff + 17),
              slice(0, n)]
    eps = 1e-6
    for sl in groups:
        v = np.zeros(n)
        v[sl] = r.standard_normal(sl.stop - sl.start)
        fd = (loss(P + eps * v) - loss(P - eps * v)) / (2 * eps)
        assert abs(fd - G @ v) <= 1e-6 * abs(G @ v), (sl, fd, G @ v)
    # padded input channels and padded K columns get no gradient
    w0 = G[lay[0][0]:lay[0][1]].reshape(16, lay[0][2])
    assert not w0[:, 27 * 8:].any()
    assert not w0[:, :27 * 8].reshape(16, 27, 8)[:, :, 2:].any()


def test_network_bf16_rounding_is_a_small_perturbation():
    """Rounding the stored activations and gradients to BF16 (the device's
    storage) changes the gradient by a few percent at most: the device is
    compared with the BF16 form, and this bounds how far that form is from
    the exact function."""
    P = U.to_bf16(_net_params(2).astype(np.float32)).astype(np.float64)
    r = rng(3)
    D = 8
    f, c = r.standard_normal((D, D, D)), r.uniform(-1, 1, (D, D, D))
    t = r.standard_normal((D, D, D))
    L0, G0 = U.net_loss_grad(P, f, c, t)
    L1, G1 = U.net_loss_grad(P, f, c, t, bf16=True)
    assert abs(L1 - L0) <= 1e-2 * L0
    assert np.linalg.norm(G1 - G0) <= 0.05 * np.linalg.norm(G0)
