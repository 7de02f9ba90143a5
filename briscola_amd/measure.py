"""Counter-collection driver: runs the kernels whose HBM traffic the
benchmark reports, for rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE).

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o run -- \
      python3 -m briscola_amd.measure [spheres|channel] [phase|prolong|
                                       residual|restrict] [fp64|mixed]

Launches, in order: 5 level-0 launches of the chosen operation on the
chosen 256^3 workload, then one plain device-to-device copy of 128 MiB (the
calibration kernel: 128 MiB read, 128 MiB written).  Operations:

  phase     red-black phase k_rbgs (spheres: face coefficients, 604 MB
            algorithmic in FP64; channel: matrix-free, 198 MB), colours
            alternating as in a sweep;
  prolong   prolongation-add k_prolong from 128^3 into 256^3 (2 n w + n/8 w);
  residual  level-0 residual k_residual (FP64 with the fused norm partials);
  restrict  restriction k_restrict of the level-0 residual to 128^3.

``python3 -m briscola_amd.measure unet [n]`` instead runs 5 forward passes
of the learned preconditioner's U-Net (random weights) on an n^3 input
(default 256), for rocprofv3 --kernel-trace --stats.
"""
import sys

OPS = {"phase": 0, "residual": 1, "restrict": 2, "prolong": 3}


def main(workload="spheres", op="phase", precision="fp64", n=256, reps=5):
    import torch
    from briscola_amd import bmg, workloads
    p = (workloads.spheres(n, seed=0) if workload == "spheres"
         else workloads.channel(n, seed=0))
    ctx = bmg.Context(p.vertices, p.bc, rel_tol=1e-8, tolerance=0.0,
                      cycle=bmg.VCYCLE, n_sweeps_pre=2, n_sweeps_post=2,
                      coarse_level=1)
    ctx.set_laplacian(p.lam)
    ctx.upload(bmg.B, 0, p.b)
    ctx.prepare()
    prec = bmg.MIXED if precision == "mixed" else bmg.FP64
    ms, by = ctx.time_op(OPS[op], 0, prec, reps)
    print(f"{workload} {op} ({precision}): {ms * 1e3:.2f} us, "
          f"{by / 1e6:.1f} MB algorithmic", flush=True)
    a = torch.empty(128 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("copy: 134217728 bytes read, 134217728 bytes written", flush=True)
    ctx.close()
    return 0


def unet_forward(n=256, reps=5):
    import torch
    from briscola_amd import unet
    n = int(n)
    net = unet.Net(seed=1)
    g = torch.Generator(device="cuda:0").manual_seed(0)
    f = torch.randn((n, n, n), device="cuda:0", generator=g)
    c = torch.randn((n, n, n), device="cuda:0", generator=g)
    net.forward(f, c)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        net.forward(f, c)
    e1.record()
    torch.cuda.synchronize()
    print(f"U-Net forward {n}^3: {e0.elapsed_time(e1) / reps:.3f} ms",
          flush=True)
    net.close()
    return 0


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "unet":
        sys.exit(unet_forward(*sys.argv[2:3]))
    sys.exit(main(*sys.argv[1:4]))
