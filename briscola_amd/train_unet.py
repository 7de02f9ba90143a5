"""Train the learned preconditioner on the device.

Pairs (b, x) come from the variable-coefficient generator (density ratio
1000, seeded spheres; briscola_amd/workloads.py) solved by this library's
FP64 flexible CG + FP64 MG cycle to a relative residual of 1e-10.  Training
runs the network's own forward/backward kernels and Adam
(``bmg_unet_train_step``) on the mean-removed squared error of the scaled
solution (see briscola_amd/unet.py for the scaling).

  python -m briscola_amd.train_unet --n 64 --samples 64 --epochs 20 \
      --out weights.npy

Returns/saves the FP32 parameters; ``python -m briscola_amd.train_unet
--eval weights.npy`` reports what the trained network gains on held-out
problems (FCG iterations and residual after the network's initial guess).
"""
import argparse
import sys
import time

import numpy as np


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solver(prob, krylov, rel_tol, precision=0, max_iter=999):
    from briscola_amd import bmg
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value, rel_tol=rel_tol,
                      tolerance=0.0, max_iter=max_iter, n_sweeps_pre=2,
                      n_sweeps_post=2, cycle=bmg.VCYCLE,
                      coarse_mode=bmg.COARSE_DIRECT, coarse_level=1,
                      precision=precision, krylov=krylov)
    ctx.set_laplacian(prob.lam)
    return ctx


def make_pairs(n, samples, seed0=1000, nspheres=32):
    """Scaled training pairs (f, c, t) as float32 cuda tensors."""
    import torch
    from briscola_amd import bmg, unet, workloads
    pairs = []
    for s in range(samples):
        prob = workloads.spheres(n, seed=seed0 + s, nspheres=nspheres)
        ctx = solver(prob, bmg.KRYLOV_FCG, 1e-10)
        x, _ = ctx.solve(prob.b.copy())
        ctx.close()
        f, c, t, _, _ = unet.scaled_pair(prob, x)
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
        pairs.append((d(f), d(c), d(t)))
    return pairs


def train(n=64, samples=64, epochs=20, lr=1e-3, seed=0, pairs=None,
          verbose=True):
    """Returns (net, per-epoch mean loss, seconds for data, seconds to
    train)."""
    from briscola_amd import unet
    t0 = time.perf_counter()
    if pairs is None:
        pairs = make_pairs(n, samples)
    t1 = time.perf_counter()
    net = unet.Net(seed=seed)
    rng = np.random.default_rng(seed)
    hist = []
    for e in range(epochs):
        # cosine decay of the learning rate over the epochs
        lr_e = lr * 0.5 * (1.0 + np.cos(np.pi * e / max(1, epochs)))
        order = rng.permutation(len(pairs))
        losses = [net.train_step(*pairs[i], lr_e) for i in order]
        hist.append(float(np.mean(losses)))
        if verbose:
            log(f"[train] epoch {e}: loss {hist[-1]:.4e} (lr {lr_e:.2e})")
    t2 = time.perf_counter()
    return net, hist, t1 - t0, t2 - t1


def evaluate(net, n=64, problems=4, seed0=9000, precision=1):
    """Per held-out problem: FCG iterations and time without and with the
    network, and the residual after its initial guess."""
    from briscola_amd import bmg, workloads
    out = []
    for s in range(problems):
        prob = workloads.spheres(n, seed=seed0 + s)
        row = {}
        for name, kry in (("fcg", bmg.KRYLOV_FCG),
                          ("fcg_learned", bmg.KRYLOV_FCG_LEARNED)):
            ctx = solver(prob, kry, 1e-8, precision)
            if kry == bmg.KRYLOV_FCG_LEARNED:
                net.attach(ctx)
            ctx.solve(prob.b.copy())   # warm-up (graph capture)
            t = time.perf_counter()
            _, st = ctx.solve(prob.b.copy())
            row[name] = dict(iters=st.iters, ms=1e3 * (time.perf_counter() - t),
                             learned_iters=st.learned_iters,
                             fallback=st.learned_fallback,
                             guess_residual=st.learned_guess_residual)
            ctx.close()
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="save parameters (.npy)")
    ap.add_argument("--eval", default=None, help="evaluate saved parameters")
    args = ap.parse_args()
    from briscola_amd import unet
    if args.eval:
        net = unet.Net()
        net.set_params(np.load(args.eval))
    else:
        net, hist, td, tt = train(args.n, args.samples, args.epochs, args.lr,
                                  args.seed)
        log(f"[train] data {td:.1f} s, training {tt:.1f} s, loss "
            f"{hist[0]:.3e} -> {hist[-1]:.3e}")
        if args.out:
            np.save(args.out, net.params())
    for row in evaluate(net, args.n):
        print(row, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
