"""Counter-collection driver: runs the kernels whose HBM traffic the
benchmark reports, for rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE).

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o run -- \
      python3 -m briscola_amd.measure [spheres|channel]

Launches, in order: 5 level-0 red-black phases of the chosen 256^3 operator
(spheres: k_rbgs<double, 2>, face coefficients, 604 MB algorithmic each;
channel: k_rbgs<double, 1>, matrix-free, 198 MB algorithmic each), colours
alternating as in a sweep, and one plain device-to-device copy of 128 MiB
(the calibration kernel: 128 MiB read, 128 MiB written).
"""
import sys


def main(workload="spheres", n=256, reps=5):
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
    ms, by = ctx.time_smoother(0, bmg.FP64, reps)
    print(f"{workload} phase: {ms * 1e3:.2f} us, {by / 1e6:.1f} MB algorithmic",
          flush=True)
    a = torch.empty(128 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("copy: 134217728 bytes read, 134217728 bytes written", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(*(sys.argv[1:2] or ["spheres"])))
