#!/usr/bin/env python
"""Benchmark: 256^3 pressure Poisson solve to 1e-8 relative residual.

Headline workload: the variable-coefficient pressure
equation div(rho^-1 grad p) = f on a 256^3 Neumann unit box with density
ratio 1000 (rho = 1000 outside, 1 inside 32 seeded spheres, interface smoothed
over 1.5 cells; f = divergence of a seeded random face velocity).  Secondary
workload: the 256^3 channel (periodic x,z, Neumann y, tanh beta=2 stretching,
Fourier-mode RHS).  Both use Briscola MG settings V(2,2) red-black
Gauss-Seidel, linear transfers, coarse level 1 (4^3) solved directly,
relTol 1e-8.  See briscola_amd/workloads.py for the generators.

On the headline workload the learned preconditioner (a 3-D U-Net trained on
the device before the timed solves, on pairs the multigrid generates at 64^3;
briscola_amd/train_unet.py) is run as initial guess and preconditioner of
flexible CG with the mixed-precision MG cycle as fall-back; its training
time is reported separately and not part of a step.

One step = one complete solve from b resident in HBM (x0 = 0) to x with
||r|| / ||r0|| <= 1e-8.  Methods: plain FP64 MG (the reference algorithm,
bitwise equal to the CPU restatement), mixed-precision MG (FP32 cycles inside
FP64 defect correction), and flexible CG preconditioned by one FP64 or mixed
MG cycle.  ``value`` is the fastest method on the headline workload; every
method is reported with iterations, final residual and its error against an
FP64 reference solution; ``equal_error`` times plain FP64 MG and FCG+mixed
again to the tolerance at which each one's error falls below 1e-8, and
``seeds`` times both on seeds 1-4 of the generator and reports the medians
over seeds 0-4.

The roofline object covers the dominant kernel of the fastest method, the
level-0 red-black smoother phase of its precision (the FP32 phase of the mixed
cycles for the mixed methods), timed live with HIP events on the library's
stream right after the timed steps of the MG solve in that precision;
``roofline_fp64`` the same for the plain FP64 solve's phase.  The CPU baseline times the CPU
restatement of the reference multigrid on a bounded sample of the same solve,
on the host cores this process may use and on one core, and scales it by the
iteration count.

The default run also reports the batched workload (``batch512``): 512
independent systems of uneven shapes and boundary conditions solved
together (solves/s), next to the same systems solved one after another.
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
MAX_ITER = 999


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solver_opts(precision, krylov=0, rel_tol=1e-8, max_iter=None):
    from briscola_amd import bmg
    return dict(rel_tol=rel_tol, tolerance=0.0,
                max_iter=MAX_ITER if max_iter is None else max_iter,
                n_sweeps_pre=2, n_sweeps_post=2, cycle=bmg.VCYCLE,
                coarse_mode=bmg.COARSE_DIRECT, coarse_level=1,
                smoother=bmg.RBGS, precision=precision, krylov=krylov)


WORKLOADS = {
    "spheres": dict(
        title="spheres 256^3",
        desc=("variable-coefficient VOF pressure div(rho^-1 grad p) = f, {n}^3 "
              "Neumann unit box, rho 1000 outside / 1 inside 32 seeded spheres "
              "(radius 0.03-0.1, interface over 1.5 cells), f = div of a random "
              "face velocity; MG V(2,2) rbgs, linear transfers, 4^3 direct "
              "coarse, relTol 1e-8"),
        kernel="k_rbgs<double, 2> (level-0 red-black phase, face-coefficient "
               "operator, 72 B per updated cell)",
        traffic="rbgs_phase_spheres256_traffic.txt",
        kernel_mixed="k_rbgs<float, 2> (level-0 red-black phase of the FP32 "
                     "cycles, FP16 face coefficients, 24 B per updated cell)",
        traffic_mixed="rbgs_phase_fp32_spheres256_traffic.txt"),
    "channel": dict(
        title="channel 256^3",
        desc=("channel {n}^3: periodic x,z, Neumann y, tanh beta=2, Fourier "
              "RHS; MG V(2,2) rbgs, linear transfers, 4^3 direct coarse, "
              "relTol 1e-8"),
        kernel="k_rbgs<double, 1> (level-0 red-black phase, interior region, "
               "matrix-free operator, 24 B per updated cell)",
        traffic="rbgs_phase_channel256_traffic.txt",
        kernel_mixed="k_rbgs<float, 1> (level-0 red-black phase of the FP32 "
                     "cycles, whole interior, matrix-free operator, 12 B per "
                     "updated cell)",
        traffic_mixed="rbgs_phase_fp32_channel256_traffic.txt"),
}


def make_problem(name, n):
    from briscola_amd import workloads
    if name == "spheres":
        return workloads.spheres(n, seed=0)
    return workloads.channel(n, seed=0)


def traffic_from_profiles(fname):
    """HBM bytes per launch of the smoother phase measured with rocprofv3
    counters (the summary committed under profiles/)."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_gpu(prob, precision, krylov, steps, warmup, profile, reps=50,
            rel_tol=1e-8, max_iter=None, net=None, deflation=0):
    import torch
    from briscola_amd import bmg
    ctx = bmg.Context(prob.vertices, prob.bc, prob.bc_value,
                      deflation=deflation,
                      **solver_opts(precision, krylov, rel_tol, max_iter))
    ctx.set_laplacian(prob.lam)
    if net is not None:
        net.attach(ctx)
    b_dev = torch.from_numpy(prob.b).to("cuda:0")
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
    # the whole V-cycle iteration (captured graph: cycle, level-0 residual
    # and norm) and the other level-0 kernels, with their algorithmic bytes
    vcyc = kern = None
    if profile:
        vcyc = ctx.time_op(bmg.TIME_ITERATION, 0, precision, 20)
        kern = {name: ctx.time_op(op, 0, precision, reps)
                for name, op in (("residual", bmg.TIME_RESIDUAL),
                                 ("restrict", bmg.TIME_RESTRICT),
                                 ("prolong", bmg.TIME_PROLONG))}
    defl = ctx.deflation_vectors()
    ctx.close()
    return dict(ms=1e3 * wall / steps, device_ms=dev_ms / steps,
                iters=stats.iters, final_residual=stats.final_residual, x=x,
                prof=prof, vcycle=vcyc, kernels=kern, deflation_vectors=defl,
                learned_iters=stats.learned_iters,
                learned_fallback=stats.learned_fallback,
                learned_guess_residual=stats.learned_guess_residual)


def run_fft(prob, steps, warmup):
    """The FFT Poisson solver (bmg_fft, the reference's FFT solver) on the
    same problem in its finite-difference form (source -b/V)."""
    import numpy as np
    import torch
    from briscola_amd import bmg
    dx, dy, dz = (np.diff(v) for v in prob.vertices)
    V = dx[:, None, None] * dy[None, :, None] * dz[None, None, :]
    b = np.zeros_like(prob.b)
    b[1:-1, 1:-1, 1:-1] = -prob.b[1:-1, 1:-1, 1:-1] / V
    fft = bmg.FFT(prob.vertices, prob.bc, prob.bc_value)
    b_dev = torch.from_numpy(b).to("cuda:0")
    x_dev = torch.zeros_like(b_dev)
    for _ in range(max(1, warmup)):
        fft.solve_device(b_dev.data_ptr(), x_dev.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fft.solve_device(b_dev.data_ptr(), x_dev.data_ptr())
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    x = x_dev.cpu().numpy()
    fft.close()
    return dict(ms=ms, x=x)


def reference_solution(name, prob):
    """FP64 reference solution and how it was obtained: the channel by the
    direct FFT (x,z) + Thomas (y) solve on the host; the variable-coefficient
    case (no direct solver at 256^3) by FP64 flexible CG with the FP64 MG
    cycle to 1e-13 on the device."""
    from oracle import direct
    if name == "channel":
        return (direct.solve_fft_thomas(prob.b, prob.vertices, bc_y=(0, 0)),
                "host FFT(x,z) + Thomas(y) direct solve, FP64")
    from briscola_amd import bmg
    r = run_gpu(prob, bmg.FP64, bmg.KRYLOV_FCG, 1, 0, False, rel_tol=1e-13,
                max_iter=3000)
    return (r["x"][1:-1, 1:-1, 1:-1],
            f"FP64 flexible CG + FP64 MG cycle to {r['final_residual']:.1e} "
            f"relative residual ({r['iters']} iterations) on the device")


def cpu_cores():
    """Host cores this process may use (the machine's share when
    OMP_NUM_THREADS states it)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_full_solve(prob):
    """The whole FP64 solve by the CPU restatement on the host's cores (once;
    validates the scaled sample)."""
    import oracle
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    opts = solver_opts(0)
    opts.pop("precision")
    opts.pop("krylov")
    o.set_opts(**opts)
    o.assemble(prob.lam)
    o.set_field(oracle.B, 0, prob.b)
    cores = cpu_cores()
    prev = oracle.get_threads()
    try:
        oracle.set_threads(cores)
        t0 = time.perf_counter()
        it, r0, r1 = o.solve()
        sec = time.perf_counter() - t0
    finally:
        oracle.set_threads(prev)
    del o
    return {"ms": round(sec * 1e3, 1), "iters": it, "cores": cores,
            "final_rel_residual": r1 / r0 if r0 else r1}


def cpu_baseline(prob, iters_fp64, sample_iters, n):
    """The CPU restatement of the reference MG on a bounded sample of the
    same FP64 solve, on one core and on the host's P cores (OpenMP over the
    planes of every per-cell loop; the same bits on any core count)."""
    import oracle
    o = oracle.Oracle(prob.vertices, prob.bc, prob.bc_value)
    opts = solver_opts(0)
    opts.pop("precision")
    opts.pop("krylov")
    o.set_opts(**opts)
    o.assemble(prob.lam)
    o.set_field(oracle.B, 0, prob.b)
    o.pin_and_eliminate()
    o.prepare_coarse()
    o.residual(0, False)
    cores = cpu_cores()
    per = {}
    prev = oracle.get_threads()
    try:
        for nt in sorted({1, cores}):
            oracle.set_threads(nt)
            t0 = time.perf_counter()
            for _ in range(sample_iters):
                o.cycle()
                o.residual(0, True)
                o.sum_sq_residual()
            per[nt] = (time.perf_counter() - t0) / sample_iters
    finally:
        oracle.set_threads(prev)
    del o
    host = platform.processor() or platform.machine()
    return {
        "value": round(per[cores] * iters_fp64 * 1e3, 1),
        "unit": "ms",
        "cores": cores,
        "kind": "port",
        "sample": (f"{sample_iters} V(2,2) iterations (cycle + residual + "
                   f"norm) of the same {n}^3 FP64 solve, timed on {cores} "
                   f"host cores ({per[cores] * 1e3:.0f} ms each) and on one "
                   f"core ({per[1] * 1e3:.0f} ms each), scaled by the "
                   f"{iters_fp64} iterations the solve takes; host {host}"),
        "one_core_ms": round(per[1] * iters_fp64 * 1e3, 1),
    }


def roofline(result, kernel, traffic):
    if not result or not result.get("prof"):
        return None
    avg_ms, per_launch = result["prof"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "kernel": kernel,
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "bytes_per_launch": per_launch}


def vcycle_roofline(result, precision_name):
    """One whole MG iteration (V(2,2) cycle over every level, level-0
    residual and norm) replayed from its captured graph between HIP events:
    algorithmic bytes of all its launches (each array a launch touches
    counted once; red-black phases as executed, i.e. 10 per level for
    V(2,2), the repeated colour of each reference sweep skipped) / device
    time.  ``frac_reference_phases`` counts the bytes of the reference's
    12 phases per level (3 per sweep, SURVEY-style p = 3) instead."""
    if not result or not result.get("vcycle"):
        return None
    ms, by = result["vcycle"]
    achieved = by / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4),
           "bytes_per_iteration": round(by),
           "iteration_us": round(ms * 1e3, 1),
           "precision": precision_name,
           "solve_device_us_per_iteration": round(
               1e3 * result["device_ms"] / max(1, result["iters"]), 1)}
    if result.get("prof"):
        # two extra level-0 phases per V(2,2) cycle (one per smoothing) in
        # the reference's sweep order; the level-0 phase dominates the bytes
        _, phase_bytes = result["prof"]
        extra = 2.0 * phase_bytes * 8.0 / 7.0
        out["frac_reference_phases"] = round(
            (by + extra) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def kernel_rooflines(result):
    """Level-0 residual (with the norm partials in FP64), restriction and
    prolongation, each launched back to back between HIP events."""
    if not result or not result.get("kernels"):
        return None
    out = {}
    for name, (ms, by) in result["kernels"].items():
        a = by / (ms * 1e-3) / 1e9
        out[name] = {"avg_launch_us": round(ms * 1e3, 2),
                     "bytes_per_launch": round(by),
                     "achieved": round(a, 1),
                     "frac": round(a / HBM_PEAK_GBS, 4)}
    return out


TARGET_ERROR = 1e-8


def _spheres_problem(args):
    n, seed = args
    from briscola_amd import workloads
    return workloads.spheres(n, seed=seed)


# Problems of the extra headline seeds, generated by worker processes that
# are started before this process touches the GPU (see main).
SEED_JOBS = {}


def start_seed_jobs(args):
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    seeds = list(range(1, args.seeds))
    if not seeds:
        return None
    pool = ProcessPoolExecutor(max_workers=min(4, len(seeds)),
                               mp_context=multiprocessing.get_context("fork"))
    for s in seeds:
        SEED_JOBS[s] = pool.submit(_spheres_problem, (args.n, s))
    return pool


def seed_solves(results, args):
    """Plain FP64 MG and FCG+mixed on the headline generator with seeds
    1..S-1 (seed 0 is the main measurement): one timed solve each (FCG+mixed
    after one untimed solve), and the medians over seeds 0..S-1."""
    import statistics
    from briscola_amd import bmg
    rows = {0: {m: (round(results[m]["ms"], 3), results[m]["iters"])
                for m in ("fp64", "fcg_mixed")}}
    for seed in sorted(SEED_JOBS):
        prob = SEED_JOBS.pop(seed).result()
        row = {}
        for mname, prec, kry, warm in (("fp64", bmg.FP64, bmg.KRYLOV_NONE, 0),
                                       ("fcg_mixed", bmg.MIXED,
                                        bmg.KRYLOV_FCG, 1)):
            r = run_gpu(prob, prec, kry, 1, warm, False)
            row[mname] = (round(r["ms"], 3), r["iters"])
        rows[seed] = row
        del prob
        log(f"[bench] spheres seed {seed}: fp64 {row['fp64'][0]:.1f} ms "
            f"({row['fp64'][1]} it), fcg_mixed {row['fcg_mixed'][0]:.1f} ms "
            f"({row['fcg_mixed'][1]} it)")
    med = {m: statistics.median(rows[s][m][0] for s in rows)
           for m in ("fp64", "fcg_mixed")}
    return {"seeds": sorted(rows),
            "ms_iters": {str(s): rows[s] for s in sorted(rows)},
            "median_ms": med,
            "median_speedup_fcg_mixed_vs_fp64": round(
                med["fp64"] / med["fcg_mixed"], 3)}


def equal_error_solves(prob, results, xe, args, net=None):
    """Plain FP64 MG and FCG+mixed solved to the relative residual that
    brings each one's error against the FP64 reference below 1e-8: the
    error at relTol 1e-8 is set by the operator's conditioning, and falls in
    proportion to the tolerance, so the tolerance is the largest power of ten
    below relTol * (1e-8 / error at relTol 1e-8), lowered once more if the
    error still misses.  One untimed and one timed solve each."""
    import math
    from briscola_amd import bmg
    from oracle import direct
    out = {"target_rel_error": TARGET_ERROR, "methods": {}}
    plan = [("fp64", bmg.FP64, bmg.KRYLOV_NONE),
            ("fcg_mixed", bmg.MIXED, bmg.KRYLOV_FCG)]
    if "fcg_deflated_mixed" in results:
        plan.append(("fcg_deflated_mixed", bmg.MIXED, bmg.KRYLOV_FCG))
    if net is not None and "fcg_learned_mixed" in results:
        plan.append(("fcg_learned_mixed", bmg.MIXED, bmg.KRYLOV_FCG_LEARNED))
    for mname, prec, kry in plan:
        err0 = results[mname]["err_vs_fp64_reference"]
        tol = 10.0 ** math.floor(math.log10(1e-8 * TARGET_ERROR / err0))
        for _ in range(2):
            log(f"[bench] spheres {mname} to relTol {tol:.0e} (equal error)")
            r = run_gpu(prob, prec, kry, 1, 1, False, rel_tol=tol,
                        max_iter=4000,
                        net=net if kry == bmg.KRYLOV_FCG_LEARNED else None,
                        deflation=int("deflated" in mname))
            err = direct.rel_error(r["x"][1:-1, 1:-1, 1:-1], xe, True)
            if err <= TARGET_ERROR:
                break
            tol /= 10.0
        out["methods"][mname] = {"rel_tol": tol, "ms": round(r["ms"], 3),
                                 "iters": r["iters"],
                                 "final_rel_residual": r["final_residual"],
                                 "err_vs_fp64_reference": err}
        log(f"[bench] spheres {mname} relTol {tol:.0e}: {r['ms']:.1f} ms, "
            f"{r['iters']} iterations, error {err:.2e}")
    m = out["methods"]
    out["speedup_fcg_mixed_vs_fp64"] = round(m["fp64"]["ms"] /
                                             m["fcg_mixed"]["ms"], 3)
    if "fcg_learned_mixed" in m:
        out["speedup_fcg_learned_mixed_vs_fp64"] = round(
            m["fp64"]["ms"] / m["fcg_learned_mixed"]["ms"], 3)
    return out


def run_workload(name, args):
    import numpy as np
    from briscola_amd import bmg
    from oracle import direct
    w = WORKLOADS[name]
    log(f"[bench] {name}: generating {args.n}^3")
    prob = make_problem(name, args.n)
    plan = [("fp64", bmg.FP64, bmg.KRYLOV_NONE),
            ("mixed", bmg.MIXED, bmg.KRYLOV_NONE),
            ("fcg_fp64", bmg.FP64, bmg.KRYLOV_FCG),
            ("fcg_mixed", bmg.MIXED, bmg.KRYLOV_FCG)]
    if name == "spheres":
        plan.append(("fcg_deflated_mixed", bmg.MIXED, bmg.KRYLOV_FCG))
    learned = None
    if name == "spheres" and not args.no_learned:
        plan.append(("fcg_learned_mixed", bmg.MIXED, bmg.KRYLOV_FCG_LEARNED))
    if args.only:
        plan = [p for p in plan if p[0] == args.only]
    results = {}
    net = None
    if any(p[2] == bmg.KRYLOV_FCG_LEARNED for p in plan):
        from briscola_amd import train_unet
        log(f"[bench] training the learned preconditioner "
            f"({args.train_n}^3, {args.train_samples} pairs, "
            f"{args.train_epochs} epochs)")
        net, hist, t_data, t_train = train_unet.train(
            args.train_n, args.train_samples, args.train_epochs, lr=2e-3,
            verbose=False)
        learned = {"train_grid": f"{args.train_n}^3",
                   "train_pairs": args.train_samples,
                   "epochs": args.train_epochs,
                   "data_s": round(t_data, 2), "train_s": round(t_train, 2),
                   "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                   "network": ("4-level 3-D U-Net 16/32/64/128 channels, 3^3 "
                               "convolutions, BF16 weights/activations, FP32 "
                               "accumulate")}
    for mname, prec, kry in plan:
        log(f"[bench] {name} {mname}: warmup {args.warmup}, steps {args.steps}")
        results[mname] = run_gpu(prob, prec, kry, args.steps, args.warmup,
                                 profile=(kry == bmg.KRYLOV_NONE),
                                 net=net if kry == bmg.KRYLOV_FCG_LEARNED
                                 else None,
                                 deflation=int("deflated" in mname))
        r = results[mname]
        log(f"[bench] {name} {mname}: {r['ms']:.2f} ms/solve, {r['iters']} "
            f"iterations, final rel. residual {r['final_residual']:.3e}")
    ref_note = None
    if not args.no_check:
        log(f"[bench] {name}: FP64 reference solution")
        xe, ref_note = reference_solution(name, prob)
        for r in results.values():
            r["err_vs_fp64_reference"] = direct.rel_error(
                r["x"][1:-1, 1:-1, 1:-1], xe, True)
    seeds = None
    if (name == "spheres" and SEED_JOBS and "fp64" in results
            and "fcg_mixed" in results):
        seeds = seed_solves(results, args)
    equal = None
    if (name == "spheres" and not args.no_check and not args.no_equal_error
            and "fp64" in results and "fcg_mixed" in results):
        equal = equal_error_solves(prob, results, xe, args, net)
    fft = None
    if name == "channel":
        log(f"[bench] {name}: FFT Poisson solver")
        fft = run_fft(prob, args.steps, args.warmup)
        log(f"[bench] {name} fft: {fft['ms']:.2f} ms/solve")
        if not args.no_check:
            fft["err_vs_fp64_reference"] = direct.rel_error(
                fft["x"][1:-1, 1:-1, 1:-1], xe, True)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] {name}: CPU baseline sample")
        iters = results.get("fp64", next(iter(results.values())))["iters"]
        cpu = cpu_baseline(prob, iters, args.cpu_sample_iters, args.n)
        if args.cpu_full:
            log(f"[bench] {name}: whole CPU solve on {cpu['cores']} cores")
            full = cpu_full_solve(prob)
            cpu["full_solve"] = full
            cpu["full_solve_vs_scaled_sample"] = round(
                full["ms"] / cpu["value"], 3)
            log(f"[bench] {name}: whole CPU solve {full['ms'] / 1e3:.1f} s, "
                f"{full['iters']} iterations (scaled sample "
                f"{cpu['value'] / 1e3:.1f} s)")
    best = min(results, key=lambda k: results[k]["ms"])
    out = {
        "workload": w["desc"].format(n=args.n),
        "best": best,
        # the dominant kernel of the fastest method: the level-0 phase of
        # its precision (FP32 cycles for the mixed methods)
        "roofline": (roofline(results.get("mixed"), w["kernel_mixed"],
                              traffic_from_profiles(w["traffic_mixed"]))
                     if "mixed" in best and "mixed" in results else
                     roofline(results.get("fp64"), w["kernel"],
                              traffic_from_profiles(w["traffic"]))),
        "roofline_fp64": roofline(results.get("fp64"), w["kernel"],
                                  traffic_from_profiles(w["traffic"])),
        "vcycle": vcycle_roofline(results.get("fp64"), "fp64"),
        "vcycle_mixed": vcycle_roofline(results.get("mixed"), "mixed"),
        "level0_kernels": kernel_rooflines(results.get("fp64")),
        "cpu_baseline": cpu,
        "fp64_reference": ref_note,
        "solves": {k: {"ms": round(v["ms"], 3),
                       "device_ms": round(v["device_ms"], 3),
                       "iters": v["iters"],
                       "deflation_vectors": v["deflation_vectors"],
                       "final_rel_residual": v["final_residual"],
                       "err_vs_fp64_reference": v.get("err_vs_fp64_reference")}
                   for k, v in results.items()},
    }
    if "fp64" in results:
        out["speedup_vs_plain_fp64_mg"] = {
            k: round(results["fp64"]["ms"] / v["ms"], 3)
            for k, v in results.items()}
    if equal is not None:
        out["equal_error"] = equal
    if seeds is not None:
        out["seeds"] = seeds
    if fft is not None:
        out["fft_solver"] = {
            "method": ("FFT(x,z) + tridiagonal(y) direct solve of the "
                       "finite-difference system (the reference's FFT "
                       "solver; rocFFT + batched Thomas)"),
            "ms": round(fft["ms"], 3),
            "err_vs_fp64_reference": fft.get("err_vs_fp64_reference")}
    if learned is not None and "fcg_learned_mixed" in results:
        r = results["fcg_learned_mixed"]
        learned.update({
            "learned_iterations": r["learned_iters"],
            "fallback_after_iteration": r["learned_fallback"],
            "residual_after_guess": r["learned_guess_residual"]})
        if "fcg_mixed" in results:
            learned["time_vs_fcg_mixed"] = round(
                r["ms"] / results["fcg_mixed"]["ms"], 3)
        out["learned"] = learned
    return out, results


def run_batch(args):
    """Batched uneven solves: ``args.batch_systems`` independent unit-box
    Poisson systems (shapes 32^3..128^3 and 64x32x128, per-axis Neumann /
    Dirichlet / periodic faces, Fourier-mode sources;
    briscola_amd/workloads.py batch_system) solved together to 1e-8 with
    mixed-precision MG, one launch per kernel over all systems per
    iteration (bmg_solve_batch), once with MG iterations and once with
    flexible CG preconditioned by the same cycle.  Also: the same systems
    solved one after another (each with its own captured iteration), and
    the errors against FP64 solutions to 1e-12."""
    import numpy as np
    import torch
    from briscola_amd import bmg, workloads
    from oracle import direct
    n = args.batch_systems
    log(f"[bench] batch: generating {n} systems")
    probs = [workloads.batch_system(i, seed=0) for i in range(n)]
    cells = sum(int(np.prod(p.n)) for p in probs)
    opts = dict(solver_opts(bmg.MIXED), coarse_level=0)
    ctxs = []
    for p in probs:
        c = bmg.Context(p.vertices, p.bc, p.bc_value, **opts)
        c.set_laplacian(None)
        ctxs.append(c)
    b0 = [torch.from_numpy(p.b).to("cuda:0") for p in probs]
    bs = [t.clone() for t in b0]
    xs = [torch.zeros_like(t) for t in b0]
    bp = [t.data_ptr() for t in bs]
    xp = [t.data_ptr() for t in xs]

    def reset():
        for b, b_0, x in zip(bs, b0, xs):
            b.copy_(b_0)
            x.zero_()
        torch.cuda.synchronize()

    batch = bmg.Batch(ctxs)
    log(f"[bench] batch: warmup {args.warmup} (set-up, plan, autotune)")
    for _ in range(max(1, args.warmup)):
        reset()
        batch.solve_device(bp, xp)
    times = []
    for _ in range(args.steps):
        reset()
        t0 = time.perf_counter()
        st = batch.solve_device(bp, xp, const_matrix=True)
        times.append(time.perf_counter() - t0)
    ms = 1e3 * sum(times) / len(times)
    launches, wg = batch.info()
    _, reshaped = batch.tuning()
    xb = [t.cpu().numpy() for t in xs]
    iters = sorted(s.iters for s in st)
    rel = max(s.final_residual / s.initial_residual for s in st
              if s.initial_residual > 0)
    log(f"[bench] batch: {ms:.1f} ms for {n} solves "
        f"({n / ms * 1e3:.0f} solves/s), iterations {iters[0]}..{iters[-1]}")
    # the same batch with flexible CG preconditioned by the same captured
    # mixed V-cycles
    for c in ctxs:
        c.set_opts(krylov=bmg.KRYLOV_FCG)
    log(f"[bench] batch fcg: warmup {args.warmup} (plan, autotune)")
    for _ in range(max(1, args.warmup)):
        reset()
        batch.solve_device(bp, xp)
    ftimes = []
    for _ in range(args.steps):
        reset()
        t0 = time.perf_counter()
        fst = batch.solve_device(bp, xp, const_matrix=True)
        ftimes.append(time.perf_counter() - t0)
    fms = 1e3 * sum(ftimes) / len(ftimes)
    flaunches, _ = batch.info()
    _, freshaped = batch.tuning()
    xf = [t.cpu().numpy() for t in xs]
    fiters = sorted(s_.iters for s_ in fst)
    frel = max(s_.final_residual / s_.initial_residual for s_ in fst
               if s_.initial_residual > 0)
    log(f"[bench] batch fcg: {fms:.1f} ms for {n} solves "
        f"({n / fms * 1e3:.0f} solves/s), iterations {fiters[0]}..{fiters[-1]}")
    for c in ctxs:
        c.set_opts(krylov=bmg.KRYLOV_NONE)
    # the first args.batch_seq systems one after another (each with its
    # own captured iteration), against a batch of the same systems
    m = min(n, args.batch_seq)
    log(f"[bench] batch: {m} systems one after another (warm-up pass)")
    for k, (ctx, b, x) in enumerate(zip(ctxs[:m], bp, xp)):
        ctx.solve_device(b, x, const_matrix=True)
        if (k + 1) % 16 == 0:
            log(f"[bench] batch: warm-up {k + 1}/{m}")
    reset()
    t0 = time.perf_counter()
    for ctx, b, x in zip(ctxs[:m], bp, xp):
        ctx.solve_device(b, x, const_matrix=True)
    torch.cuda.synchronize()
    seq_ms = 1e3 * (time.perf_counter() - t0)
    equal = all(np.array_equal(a.view(np.int64), t.cpu().numpy().view(np.int64))
                for a, t in zip(xb[:m], xs[:m]))
    sub = bmg.Batch(ctxs[:m])
    reset()
    sub.solve_device(bp[:m], xp[:m], const_matrix=True)
    reset()
    t0 = time.perf_counter()
    sub.solve_device(bp[:m], xp[:m], const_matrix=True)
    sub_ms = 1e3 * (time.perf_counter() - t0)
    sub.close()
    log(f"[bench] batch: {m} systems one after another {seq_ms:.1f} ms, "
        f"batched {sub_ms:.1f} ms; bitwise equal to the batch: {equal}")
    err = ferr = None
    if not args.no_check:
        for c in ctxs:
            c.set_opts(precision=bmg.FP64, rel_tol=1e-12)
        reset()
        batch.solve_device(bp, xp)
        err = max(direct.rel_error(a[1:-1, 1:-1, 1:-1],
                                   t.cpu().numpy()[1:-1, 1:-1, 1:-1],
                                   bmg.DIRICHLET not in p.bc)
                  for a, t, p in zip(xb, xs, probs))
        ferr = max(direct.rel_error(a[1:-1, 1:-1, 1:-1],
                                    t.cpu().numpy()[1:-1, 1:-1, 1:-1],
                                    bmg.DIRICHLET not in p.bc)
                   for a, t, p in zip(xf, xs, probs))
    batch.close()
    for c in ctxs:
        c.close()
    shapes = {}
    for p in probs:
        k = "x".join(str(q) for q in p.n)
        shapes[k] = shapes.get(k, 0) + 1
    return {
        "workload": (f"{n} independent unit-box Poisson systems (seed 0), "
                     f"{cells / 1e6:.1f} M cells in total, shapes {shapes}, "
                     "per-axis Neumann/Dirichlet/periodic faces, Fourier-mode "
                     "sources; mixed-precision MG V(2,2) rbgs, linear "
                     "transfers, direct coarsest level, relTol 1e-8; as MG "
                     "iterations and as flexible CG preconditioned by the "
                     "same cycle"),
        "metric": "solves/s to 1e-8 (whole batch, inputs resident in HBM)",
        "method": "fcg_mixed" if fms < ms else "mg_mixed",
        "solves_per_s": round(n / min(ms, fms) * 1e3, 1),
        "ms_per_batch": round(min(ms, fms), 2),
        "mg_mixed": {"solves_per_s": round(n / ms * 1e3, 1),
                     "ms_per_batch": round(ms, 2),
                     "iterations_min_median_max": [iters[0],
                                                   iters[len(iters) // 2],
                                                   iters[-1]],
                     "max_final_rel_residual": rel,
                     "max_err_vs_fp64_1e-12": err,
                     "launches_per_iteration": launches,
                     "launches_autotuned_to_other_tile_shape": reshaped},
        "fcg_mixed": {"solves_per_s": round(n / fms * 1e3, 1),
                      "ms_per_batch": round(fms, 2),
                      "iterations_min_median_max": [fiters[0],
                                                    fiters[len(fiters) // 2],
                                                    fiters[-1]],
                      "max_final_rel_residual": frel,
                      "max_err_vs_fp64_1e-12": ferr,
                      "launches_per_iteration": flaunches,
                      "launches_autotuned_to_other_tile_shape": freshaped},
        "subset": {"systems": m,
                   "one_after_another_ms": round(seq_ms, 2),
                   "batched_ms": round(sub_ms, 2),
                   "batched_speedup": round(seq_ms / sub_ms, 3),
                   "bitwise_equal": equal},
        "iterations_min_median_max": [iters[0], iters[len(iters) // 2],
                                      iters[-1]],
        "max_final_rel_residual": rel,
        "max_err_vs_fp64_1e-12": err,
        "launches_per_iteration": launches,
        "virtual_workgroups_per_iteration": wg,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", choices=["spheres", "channel", "batch",
                                           "both", "all"],
                    default="all",
                    help="headline is spheres; 'both' adds the channel, "
                         "'all' also the batched systems")
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-seq", type=int, default=64,
                    help="systems solved one after another for comparison")
    ap.add_argument("--cpu-sample-iters", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the whole FP64 solve on the CPU (minutes)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the FP64 reference solution")
    ap.add_argument("--only", choices=["fp64", "mixed", "fcg_fp64", "fcg_mixed",
                                       "fcg_deflated_mixed",
                                       "fcg_learned_mixed"], default=None)
    ap.add_argument("--seeds", type=int, default=5,
                    help="headline seeds 0..S-1 for the median (1: seed 0 only)")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the solves to a 1e-8 error on the headline")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned preconditioner")
    ap.add_argument("--train-n", type=int, default=64)
    ap.add_argument("--train-samples", type=int, default=32)
    ap.add_argument("--train-epochs", type=int, default=12)
    ap.add_argument("--max-iter", type=int, default=999,
                    help="cap on MG iterations (profiling runs only)")
    args = ap.parse_args()
    global MAX_ITER
    MAX_ITER = args.max_iter
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(each MI355X runs one brick; see the sharding design).",
              file=sys.stderr)
        return 2

    # workers for the extra seeds fork now, before any GPU work
    pool = None
    if (args.workload in ("spheres", "both", "all") and args.seeds > 1
            and args.only is None):
        pool = start_seed_jobs(args)

    if args.workload == "batch":
        out = run_batch(args)
        print(json.dumps({"batch": out}), flush=True)
        return 0
    head = "spheres" if args.workload in ("spheres", "both", "all") \
        else "channel"
    main_out, main_res = run_workload(head, args)
    extra = None
    if args.workload in ("both", "all"):
        extra, _ = run_workload("channel", args)
    batch = run_batch(args) if args.workload == "all" else None

    best = main_out["best"]
    rb = main_res[best]
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
        "dtype": ("f64" if best in ("fp64", "fcg_fp64") else
                  "bf16+f32+f64" if best == "fcg_learned_mixed" else "f32+f64"),
        "data": "synthetic",
        "config": {"workload": main_out["workload"], "method": best},
        "roofline": main_out["roofline"],
        "roofline_fp64": main_out["roofline_fp64"],
        "vcycle": main_out["vcycle"],
        "vcycle_mixed": main_out["vcycle_mixed"],
        "level0_kernels": main_out["level0_kernels"],
        "cpu_baseline": main_out["cpu_baseline"],
        "fp64_reference": main_out["fp64_reference"],
        "solves": main_out["solves"],
    }
    if "speedup_vs_plain_fp64_mg" in main_out:
        line["speedup_vs_plain_fp64_mg"] = main_out["speedup_vs_plain_fp64_mg"]
    if "equal_error" in main_out:
        line["equal_error"] = main_out["equal_error"]
        # the metric's target as one figure per method: relative residual
        # <= 1e-8 AND error against the FP64 reference <= 1e-8
        tt = {}
        for m, v in main_out["solves"].items():
            e = main_out["equal_error"]["methods"].get(m)
            err = v.get("err_vs_fp64_reference")
            if err is not None and err <= TARGET_ERROR:
                tt[m] = {"ms": v["ms"], "rel_tol": 1e-8,
                         "err_vs_fp64_reference": err}
            elif e is not None:
                tt[m] = {"ms": e["ms"], "rel_tol": e["rel_tol"],
                         "err_vs_fp64_reference": e["err_vs_fp64_reference"]}
        if "fp64" in tt:
            for m in tt:
                tt[m]["speedup_vs_fp64"] = round(tt["fp64"]["ms"] /
                                                 tt[m]["ms"], 3)
        line["time_to_residual_and_error_1e-8"] = tt
    if "seeds" in main_out:
        line["seeds"] = main_out["seeds"]
    if "learned" in main_out:
        line["learned"] = main_out["learned"]
    if extra is not None:
        line["channel256"] = extra
    if batch is not None:
        line["batch512"] = batch
    if pool is not None:
        pool.shutdown(wait=True, cancel_futures=True)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
