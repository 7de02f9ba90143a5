#!/usr/bin/env python
"""Benchmark: 256^3 channel-flow pressure Poisson solve to 1e-8 rel-res.

Workload: periodic in x and z, Neumann in y, tanh (beta = 2) wall-normal
stretching, Fourier-mode right-hand side (see briscola_amd/workloads.py),
Briscola MG settings V(2,2) red-black Gauss-Seidel, linear transfers, 4^3
coarse level solved directly, relTol 1e-8.  One step = one complete solve
from b resident in HBM (x0 = 0) to x with ||r|| / ||r0|| <= 1e-8, run both in
FP64 (the reference arithmetic, bitwise equal to the CPU restatement) and in
mixed precision (FP32 cycles inside FP64 defect correction).  ``value`` is the
faster of the two; both are reported.

The roofline object covers the dominant kernel of the FP64 solve, the
level-0 red-black smoother phase, timed live with HIP events on the
library's stream over the timed steps.  The CPU baseline times the CPU
restatement of the reference multigrid (one core) on a bounded sample of the
same solve and scales it by the iteration count; the same leg checks each GPU
solution against an FP64 direct solve (FFT in x,z + Thomas in y).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("256³ Poisson time to 1e-8 rel-res: learned/mixed vs plain FP64 MG; "
          "HBM GB/s")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


MAX_ITER = 999


def solver_opts(precision, krylov=0):
    from briscola_amd import bmg
    return dict(rel_tol=1e-8, tolerance=0.0, max_iter=MAX_ITER, n_sweeps_pre=2,
                n_sweeps_post=2, cycle=bmg.VCYCLE, coarse_mode=bmg.COARSE_DIRECT,
                coarse_level=1, smoother=bmg.RBGS, precision=precision,
                krylov=krylov)


def traffic_from_profiles():
    """HBM bytes per launch of the smoother phase measured with rocprofv3
    counters (written by hand into profiles/ after the PMC runs)."""
    path = os.path.join(ROOT, "profiles", "rbgs_phase_channel256_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_gpu(prob, precision, krylov, steps, warmup, profile, reps=50):
    import torch
    from briscola_amd import bmg
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      **solver_opts(precision, krylov))
    ctx.set_laplacian(prob.lam)
    b_host = torch.from_numpy(prob.b)
    b_dev = b_host.to("cuda:0")
    x_dev = torch.zeros_like(b_dev)
    stats = None
    for _ in range(warmup):
        x_dev.zero_()
        torch.cuda.synchronize()
        ctx.solve_device(b_dev.data_ptr(), x_dev.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(steps):
        x_dev.zero_()
        stats = ctx.solve_device(b_dev.data_ptr(), x_dev.data_ptr())
        dev_ms += stats.solve_ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    x = x_dev.cpu().numpy()
    # dominant kernel: the level-0 red-black phase, timed with HIP events on
    # the solver's stream right after the timed steps (overwrites x on the
    # device, which was copied out above)
    prof = ctx.time_smoother(0, precision, reps) if profile else None
    ctx.close()
    return dict(ms=1e3 * wall / steps, device_ms=dev_ms / steps, iters=stats.iters,
                final_residual=stats.final_residual, x=x, prof=prof)


def cpu_baseline_and_check(prob, results, iters_fp64, sample_iters):
    """CPU restatement of the reference MG on a bounded sample, and the FP64
    direct-solve error of each GPU solution."""
    import numpy as np
    import oracle
    from oracle import direct
    out = {}
    xe = direct.solve_fft_thomas(prob.b, prob.vertices, bc_y=(0, 0))
    for name, r in results.items():
        r["err_vs_fp64_oracle"] = direct.rel_error(r["x"][1:-1, 1:-1, 1:-1], xe, True)
    o = oracle.Oracle(prob.vertices, prob.bc)
    opts = solver_opts(0)
    opts.pop("precision")
    opts.pop("krylov")
    o.set_opts(**opts)
    o.assemble()
    o.set_field(oracle.B, 0, prob.b)
    o.pin_and_eliminate()
    o.prepare_coarse()
    o.residual(0, False)
    t0 = time.perf_counter()
    for _ in range(sample_iters):
        o.cycle()
        o.residual(0, True)
        o.sum_sq_residual()
    per_iter = (time.perf_counter() - t0) / sample_iters
    out["cpu"] = {
        "value": round(per_iter * iters_fp64 * 1e3, 1),
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": (f"{sample_iters} V(2,2) iterations (cycle + residual + norm) of "
                   f"the same 256^3 FP64 solve on one host core "
                   f"({per_iter * 1e3:.0f} ms each), scaled by the "
                   f"{iters_fp64} iterations the solve takes; "
                   f"host {platform.processor() or platform.machine()}"),
    }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-sample-iters", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the FP64 direct check")
    ap.add_argument("--only", choices=["fp64", "mixed", "fcg_fp64", "fcg_mixed"],
                    default=None)
    ap.add_argument("--max-iter", type=int, default=999,
                    help="cap on MG iterations (profiling runs only)")
    args = ap.parse_args()
    global MAX_ITER
    MAX_ITER = args.max_iter
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(each MI355X runs one brick; see the sharding design).", file=sys.stderr)
        return 2

    from briscola_amd import bmg, workloads
    log(f"[bench] generating channel {args.n}^3")
    prob = workloads.channel(args.n, seed=0)
    results = {}
    plan = [("fp64", bmg.FP64, bmg.KRYLOV_NONE), ("mixed", bmg.MIXED, bmg.KRYLOV_NONE),
            ("fcg_fp64", bmg.FP64, bmg.KRYLOV_FCG), ("fcg_mixed", bmg.MIXED, bmg.KRYLOV_FCG)]
    if args.only:
        plan = [p for p in plan if p[0] == args.only]
    for name, prec, kry in plan:
        log(f"[bench] {name}: warmup {args.warmup}, steps {args.steps}")
        results[name] = run_gpu(prob, prec, kry, args.steps, args.warmup,
                                profile=(kry == bmg.KRYLOV_NONE))
        r = results[name]
        log(f"[bench] {name}: {r['ms']:.2f} ms/solve, {r['iters']} iterations, "
            f"final rel. residual {r['final_residual']:.3e}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline sample and FP64 direct check")
        iters = results.get("fp64", next(iter(results.values())))["iters"]
        cpu = cpu_baseline_and_check(prob, results, iters,
                                     args.cpu_sample_iters)["cpu"]

    best = min(results, key=lambda k: results[k]["ms"])
    rb = results[best]
    roof = None
    if "fp64" in results and results["fp64"]["prof"]:
        avg_ms, per_launch = results["fp64"]["prof"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(),
                "kernel": "k_rbgs<double, 1> (level-0 red-black phase, interior region)",
                "avg_launch_us": round(avg_ms * 1e3, 2),
                "bytes_per_launch": per_launch}
    line = {
        "metric": METRIC,
        "value": round(rb["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rb["ms"], 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if best in ("fp64", "fcg_fp64") else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"channel {args.n}^3: periodic x,z, Neumann y, tanh "
                                f"beta=2, Fourier RHS; MG V(2,2) rbgs, linear "
                                f"transfers, 4^3 direct coarse, relTol 1e-8"),
                   "method": best},
        "roofline": roof,
        "cpu_baseline": cpu,
        "solves": {k: {"ms": round(v["ms"], 3), "device_ms": round(v["device_ms"], 3),
                       "iters": v["iters"],
                       "final_rel_residual": v["final_residual"],
                       "err_vs_fp64_oracle": v.get("err_vs_fp64_oracle")}
                   for k, v in results.items()},
    }
    if "fp64" in results:
        line["speedup_vs_plain_fp64_mg"] = {
            k: round(results["fp64"]["ms"] / v["ms"], 3) for k, v in results.items()}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
