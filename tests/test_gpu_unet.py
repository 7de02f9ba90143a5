"""Learned preconditioner on the device: the BF16 matrix-core convolution
kernels against the CPU reference (oracle/unet_ref.py) -- forward, data
gradient and weight gradient --, a training step that lowers the loss, and
the learned flexible-CG solve, which must reach the same tolerance and the
same solution as the plain FP64 solve (the fall-back to MG cycles keeps it
convergent whatever the network predicts)."""
import numpy as np
import pytest

from oracle import unet_ref as U
from briscola_amd import bmg, unet, workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev_bf16(a):
    """float32 array of BF16 values -> bfloat16 cuda tensor (same bits)."""
    t = torch.from_numpy(U.bf16_bits(a).view(np.int16)).to("cuda:0")
    return t.view(torch.bfloat16)


def host(t):
    return t.float().cpu().numpy()


def close(gpu, ref, rel=2e-2):
    scale = np.abs(ref).max() + 1e-30
    err = np.abs(gpu - ref).max() / scale
    assert err < rel, f"max error {err:.3e} of the largest value"


@pytest.mark.parametrize("shape,cin,cout,relu", [
    ((8, 8, 8), 16, 16, True),
    ((8, 16, 8), 8, 32, False),
    ((16, 8, 8), 48, 48, True),
    ((8, 8, 16), 192, 64, True),
    # partial output tiles (x, y and z extents not multiples of the tile)
    ((6, 10, 20), 32, 96, True),
    ((4, 6, 12), 96, 192, False),
    ((3, 5, 7), 8, 128, True),
])
def test_conv_forward(gpu, shape, cin, cout, relu):
    r = np.random.default_rng(cin + cout)
    x = U.to_bf16(r.standard_normal(shape + (cin,)).astype(np.float32))
    kp = unet.round32(27 * cin)
    w = np.zeros((cout, kp), np.float32)
    w[:, :27 * cin] = U.to_bf16(
        (r.standard_normal((cout, 27 * cin)) / np.sqrt(27 * cin)).astype(np.float32))
    bias = r.standard_normal(cout).astype(np.float32)
    ref = U.conv(x, w[:, :27 * cin], bias, relu)
    y = unet.op_conv(dev_bf16(x), dev_bf16(w),
                     torch.from_numpy(bias).to("cuda:0"), cout, relu)
    close(host(y), ref)


@pytest.mark.parametrize("cin,cout", [(16, 32), (48, 16), (96, 32)])
def test_conv_data_gradient(gpu, cin, cout):
    r = np.random.default_rng(7 + cin)
    shape = (8, 8, 16)
    dz = U.to_bf16(r.standard_normal(shape + (cout,)).astype(np.float32))
    kp = unet.round32(27 * cin)
    w = np.zeros((cout, kp), np.float32)
    w[:, :27 * cin] = U.to_bf16(
        (r.standard_normal((cout, 27 * cin)) / np.sqrt(27 * cout)).astype(np.float32))
    ref = U.conv_dgrad(dz, w[:, :27 * cin], cin)
    dx = unet.op_conv_dgrad(dev_bf16(dz), dev_bf16(w), cin)
    close(host(dx), ref)


@pytest.mark.parametrize("cin,cout,chunk", [(8, 16, 256), (32, 64, 160),
                                            (48, 16, 1024)])
def test_conv_weight_gradient(gpu, cin, cout, chunk):
    r = np.random.default_rng(11 + cin)
    shape = (8, 16, 8)
    x = U.to_bf16(r.standard_normal(shape + (cin,)).astype(np.float32))
    dz = U.to_bf16(r.standard_normal(shape + (cout,)).astype(np.float32))
    kp = unet.round32(27 * cin)
    ref = U.conv_wgrad(x, dz, kp)
    dw = unet.op_conv_wgrad(dev_bf16(x), dev_bf16(dz), kp, chunk)
    got = dw.cpu().numpy()
    # FP32 accumulation of exact BF16 products
    assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()
    assert np.all(got[:, 27 * cin:] == 0.0)


def small_problem(n=16, seed=3):
    return W.spheres(n, seed=seed, nspheres=4)


def test_training_lowers_the_loss(gpu):
    prob = small_problem()
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=1e-10,
                      tolerance=0.0, max_iter=500, n_sweeps_pre=2,
                      n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_level=1,
                      krylov=bmg.KRYLOV_FCG)
    ctx.set_laplacian(prob.lam)
    x, _ = ctx.solve(prob.b.copy())
    f, c, t, _, _ = unet.scaled_pair(prob, x)
    net = unet.Net(seed=1)
    d = lambda a: torch.from_numpy(a).to("cuda:0")
    F, C, T = d(f), d(c), d(t)
    losses = [net.train_step(F, C, T, 1e-3) for _ in range(40)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], losses[::8]
    # the forward entry point gives the trained prediction
    p = net.forward(F, C).cpu().numpy()
    e = p - p.mean() - t
    assert np.mean(e * e) < losses[0]


@pytest.mark.parametrize("seed", [1, 4])
def test_network_gradient_matches_oracle(gpu, seed):
    """One training step's gradient (every convolution's weights and bias,
    the head; bmg_unet_get_grads) and loss against the float64 restatement
    of the whole network (oracle/unet_ref.py, pinned by central differences
    in tests/test_unet_oracle.py) with BF16 rounding where the device stores
    BF16, on a 16^3 input: the backward kernels of the ReLU, pooling,
    upsampling, concatenation, head, bias sums and loss all enter it."""
    D = 16
    r = np.random.default_rng(seed)
    f = r.standard_normal((D, D, D)).astype(np.float32)
    c = r.uniform(-1, 1, (D, D, D)).astype(np.float32)
    t = r.standard_normal((D, D, D)).astype(np.float32)
    net = unet.Net(seed=seed)
    P = net.params().astype(np.float64)
    d = lambda a: torch.from_numpy(a).to("cuda:0")
    loss = net.train_step(d(f), d(c), d(t), 0.0)
    G = net.grads().astype(np.float64)
    L, Gr = U.net_loss_grad(P, f.astype(np.float64), c.astype(np.float64),
                            t.astype(np.float64), bf16=True)
    L0, G0 = U.net_loss_grad(P, f.astype(np.float64), c.astype(np.float64),
                             t.astype(np.float64))
    lay, hoff, n = U.param_layout()
    names = ["E0", "E1", "E2", "E3", "D2", "D1", "D0"]
    rows = [("loss", abs(loss - L) / L, 1.0)]
    groups = []
    for name, (wo, bo, kp), (cin, cout, _, _) in zip(names, lay, U.CONVS):
        groups.append((name + " w", slice(wo, wo + cout * kp)))
        groups.append((name + " b", slice(bo, bo + cout)))
    groups.append(("head", slice(hoff, hoff + 17)))
    worst = 0.0
    for name, sl in groups:
        g, gr, g0 = G[sl], Gr[sl], G0[sl]
        rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
        rel0 = np.linalg.norm(g - g0) / np.linalg.norm(g0)
        rows.append((name, rel, cos, rel0))
        worst = max(worst, rel)
    print(f"\n[unet grad] seed {seed}: loss device {loss:.6g} oracle {L:.6g}")
    for row in rows[1:]:
        print(f"[unet grad] {row[0]:7s} rel err vs BF16 oracle {row[1]:.2e} "
              f"cos {row[2]:.6f}; vs exact float64 {row[3]:.2e}")
    assert abs(loss - L) <= 1e-3 * L and abs(L - L0) <= 2e-2 * L0
    for name, rel, cos, rel0 in rows[1:]:
        assert rel < 0.05 and cos > 0.998, (name, rel, cos)
        assert rel0 < 0.15, (name, rel0)
    # the padded channels of the first convolution get no gradient
    w0 = G[lay[0][0]:lay[0][1]].reshape(16, lay[0][2])
    assert not w0[:, 27 * 8:].any()
    assert not w0[:, :27 * 8].reshape(16, 27, 8)[:, :, 2:].any()
    net.close()


@pytest.mark.parametrize("precision", [bmg.FP64, bmg.MIXED], ids=["fp64", "mixed"])
def test_learned_fcg_reaches_tolerance(gpu, precision):
    prob = W.spheres(32, seed=2, nspheres=6)
    base = dict(rel_tol=1e-8, tolerance=0.0, max_iter=500, n_sweeps_pre=2,
                n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_level=1,
                precision=precision)
    ref = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(base, krylov=bmg.KRYLOV_FCG, rel_tol=1e-12,
                             precision=bmg.FP64))
    ref.set_laplacian(prob.lam)
    xr, _ = ref.solve(prob.b.copy())
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **dict(base, krylov=bmg.KRYLOV_FCG_LEARNED))
    ctx.set_laplacian(prob.lam)
    net = unet.Net(seed=5)
    net.attach(ctx)
    x, st = ctx.solve(prob.b.copy())
    assert st.final_residual / st.initial_residual <= 1e-8
    xi, ri = x[1:-1, 1:-1, 1:-1], xr[1:-1, 1:-1, 1:-1]
    err = np.linalg.norm((xi - xi.mean()) - (ri - ri.mean())) / np.linalg.norm(ri - ri.mean())
    assert err < 1e-5
    assert st.learned_fallback != -1 or st.learned_iters == st.iters
    # and against an FP64 direct solve of the oracle's eliminated system
    import oracle
    from oracle import direct
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    o.assemble(prob.lam)
    o.pin_and_eliminate()
    xs = direct.solve_sparse(o.stencil(0), prob.b, prob.bc)
    err_d = direct.rel_error(xi, xs, True)
    print(f"[learned fcg] precision={precision}: {st.iters} iterations, "
          f"error vs sparse direct {err_d:.2e}", flush=True)
    assert err_d < 1e-5


def guess_error(net, prob, xr):
    """Relative L2 error (mean removed) of the network's initial guess
    x = U(b / rms(b)) * rms(b) / (h s) against the solution xr."""
    f, c, _, rms, hs = unet.scaled_pair(prob)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    u = net.forward(d(f), d(c)).cpu().numpy().astype(np.float64)
    g = (u - u.mean()) * rms / hs
    ri = xr[1:-1, 1:-1, 1:-1]
    ri = ri - ri.mean()
    return float(np.linalg.norm(g - ri) / np.linalg.norm(ri))


def test_what_training_gains(gpu, capsys):
    """What the trained network gains over the method without it, measured
    on held-out problems of its training distribution (spheres, density
    ratio 1000, 32^3; trained on 32 other seeds with the project's own
    kernels).  Training must bring the network's initial guess closer to the
    solution than the untrained network's and than the zero guess FCG starts
    from without a network.  The learned flexible-CG solve is then timed
    against FCG with the MG cycle alone and both iteration counts are
    printed: at this training budget the guess is still tens of percent off,
    the solver keeps it only when it lowers the residual, and the learned
    solve reaches the tolerance in no more iterations than the method without
    the network plus the fall-back's restart."""
    from briscola_amd import train_unet
    net, hist, _, _ = train_unet.train(32, 32, 20, lr=2e-3, verbose=False)
    raw = unet.Net(seed=7)
    assert hist[-1] < 0.2 * hist[0], hist
    rows = []
    for s in range(2):
        prob = W.spheres(32, seed=9000 + s)
        ref = train_unet.solver(prob, bmg.KRYLOV_FCG, 1e-12)
        xr, _ = ref.solve(prob.b.copy())
        ref.close()
        e_trained = guess_error(net, prob, xr)
        e_raw = guess_error(raw, prob, xr)
        # the zero guess has relative error 1
        assert e_trained < 1.0 and e_trained < e_raw, (e_trained, e_raw)
        it = {}
        for name, kry in (("fcg", bmg.KRYLOV_FCG),
                          ("fcg_learned", bmg.KRYLOV_FCG_LEARNED)):
            ctx = train_unet.solver(prob, kry, 1e-8, bmg.MIXED)
            if kry == bmg.KRYLOV_FCG_LEARNED:
                net.attach(ctx)
            _, st = ctx.solve(prob.b.copy())
            ctx.close()
            assert st.final_residual / st.initial_residual <= 1e-8
            it[name] = (st.iters, st.learned_guess_residual)
        assert it["fcg_learned"][0] <= it["fcg"][0] + 2, it
        rows.append((s, e_trained, e_raw, it))
    with capsys.disabled():
        for s, et, er, it in rows:
            print(f"\n[learned gain] held-out seed {s}: guess error trained "
                  f"{et:.3f} / untrained {er:.3g} / zero 1; FCG iterations "
                  f"{it['fcg'][0]} without the network, "
                  f"{it['fcg_learned'][0]} with it (guess residual "
                  f"{it['fcg_learned'][1]:.3g} of the zero guess's)")
