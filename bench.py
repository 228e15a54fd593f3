"""Benchmark: H-LU factor + refined solve per second on one MI355X.

Workload: N = 16384 points uniform in the unit square (Morton order),
A_ij = 1/(|p_i - p_j| + 1e-3) + (N/4) delta_ij, HODLR with 6 levels, leaf 256,
rank 32, 16 seeded N(0,1) right-hand sides.  One step = restore the factor
arena from the resident FP64 H-matrix, factor it with the persistent VM,
solve the 16 right-hand sides and refine in FP64 until the relative residual
||b - A_H x|| / ||b|| <= 1e-10.  Inputs are resident in HBM before timing.

`value` is the mixed-precision (BF16/FP32 MFMA factor + FP64 refinement)
throughput; the plain FP64 path is measured beside it in the same run.
The CPU baseline runs the same factorization and solve instruction streams
with the CPU oracle executor on the host cores, refined in FP64 to the same
tolerance (a bounded sample of about 10 s); its refined solution is the FP64
oracle solution every GPU solution's error is stated against, and the
reference's CPU configuration (n = 1024, all-dense, one right-hand side) is
timed beside it.  The line also carries the reference's run and scheduler
summaries, and two further records: the batched configuration (512 matrices
in one persistent launch) and the serving configuration (a Poisson request
stream, p50 / p99 latency at 80 % of the closed-loop peak).
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_BYTES_PER_S = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def factor_flop_split(prog):
    """(leaf getrf/trsm FLOPs, gemm/accum FLOPs) of a factor program."""
    leaf, upd = 0, 0
    for i in range(prog.info()["n_insts"]):
        op = int(prog.inst_words(i)[2])
        f = prog.inst_flops(i)
        if op in (1, 2, 3, 9):
            leaf += f
        else:
            upd += f
    return leaf, upd


def lower_bound_ms(prog, mixed):
    """T_lb = sum_i max(F_i / P_i, B_i / BW) over the factor program's
    instructions: P_i the FP64 MFMA peak (plain), or FP32 for leaf getrf /
    trsm and BF16 for products (mixed); B_i the instruction's algorithmic
    bytes at the storage width (8 bytes FP64, 4 FP32); BW = 8 TB/s."""
    t = 0.0
    for i in range(prog.info()["n_insts"]):
        op = int(prog.inst_words(i)[2])
        f = prog.inst_flops(i)
        b = prog.inst_bytes(i) * (0.5 if mixed else 1.0)
        p = (PEAK_TFLOPS["fp32"] if op in (1, 2, 3, 9) else PEAK_TFLOPS["bf16"]) if mixed else PEAK_TFLOPS["fp64"]
        t += max(f / (p * 1e12), b / (HBM_BYTES_PER_S))
    return 1e3 * t


def run_precision(pp, torch, ctx, hm, arena_dev, b, prec, steps, warmup, workers, wavefronts, nrhs, tol):
    solver = pp.Solver(ctx, hm, arena_dev.data_ptr(), precision=prec, workers=workers, wavefronts=wavefronts,
                       nrhs_max=nrhs)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    reps = []

    def step():
        solver.factor()
        rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=tol, max_iter=60)
        return rep, solver.factor_stream_ms()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = []
    lat = []
    for _ in range(steps):
        s0 = time.perf_counter()
        rep, fms = step()
        lat.append(time.perf_counter() - s0)
        reps.append(rep)
        kms.append(fms)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    x = dx.cpu().numpy().copy()
    out = dict(seconds=dt, per_step_ms=1e3 * dt / steps, factor_kernel_ms=float(np.mean(kms)),
               residual=max(r["residual"] for r in reps), iterations=max(r["iterations"] for r in reps),
               converged=all(r["converged"] for r in reps), factor_flops=solver.factor_flops,
               p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)))
    solver.close()
    return out, x


def cpu_baseline(pp, orc, hm, arena, b, threads, tol, min_seconds=10.0, max_reps=8):
    """The same factor + solve + FP64 refinement step on the host: the oracle
    executor runs the factorization and substitution-solve streams on
    `threads` workers, the residual uses a numpy H-matvec.  Repeated until
    `min_seconds` have passed (at most `max_reps` times); returns (seconds
    per step, steps, worst residual, iterations, the refined solution)."""
    prog = pp.Program.factor(hm, threads)
    info = prog.info()
    sprog = pp.Program.solve(hm, b.shape[1], threads)
    sinfo = sprog.info()
    rnd = orc.sketch(orc.sketch_need(prog.streams()))
    leaves = hm.leaves()
    bn = np.linalg.norm(b, axis=0)
    t0 = time.perf_counter()
    reps, worst, iters = 0, 0.0, 0
    while True:
        fac = arena.copy()
        tmp = np.zeros(max(info["temp_elems"], 1))
        orc.run_streams(prog.streams(), [fac, tmp, None], info["n_insts"], info["scratch_elems"], rnd=rnd,
                        threads=threads)

        def apply_inv(r):
            x = np.ascontiguousarray(r.copy())
            orc.run_streams(sprog.streams(), [fac, None, x], sinfo["n_insts"], sinfo["scratch_elems"],
                            threads=threads)
            return x

        x = apply_inv(b)
        it = 0
        while True:
            r = b - orc.hmatvec(leaves, arena, x)
            res = float(np.max(np.linalg.norm(r, axis=0) / bn))
            if res <= tol or it >= 60:
                break
            x = x + apply_inv(r)
            it += 1
        reps += 1
        worst, iters = max(worst, res), max(iters, it)
        if time.perf_counter() - t0 >= min_seconds or reps >= max_reps:
            break
    return (time.perf_counter() - t0) / reps, reps, worst, iters, x


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_mfma_busy(precision):
    """MFMA busy fraction of vm_factor_kernel from the committed PMC summary
    (mean over its dispatches), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "vm_factor_kernel_mfma_hodlr_n16384_l6_r32.txt")
    vals = []
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 7 and parts[0] == precision:
                    vals.append(float(parts[6]))
    except (OSError, ValueError):
        return None
    return sum(vals) / len(vals) if vals else None


def summary_columns(pp, hm, prog, factor_flops, factor_ms):
    """The reference's run summary (src/launcher.cu:174-212: actual and
    dense-LU FLOPs, FLOPs compression ratio, GPU and equivalent dense-LU
    rates) and scheduler summary (rts/scheduler.cpp:406-427)."""
    tr = importlib.import_module("pastel-palettes_amd.trace")
    dense = pp.dense_lu_flops(hm.n)
    gflops = factor_flops / (factor_ms * 1e-3) / 1e9
    return {"actual_flops": int(factor_flops), "dense_lu_flops": int(dense),
            "flops_compression_pct": 100.0 * factor_flops / dense,
            "gpu_gflops": gflops, "equivalent_dense_lu_tflops": dense / (factor_ms * 1e-3) / 1e12,
            "scheduler": tr.schedule_summary(prog)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--rank", type=int, default=32)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--wavefronts", type=int, default=4)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 256 MiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--calibrate-mfma", action="store_true",
                    help="run only three FP64 and three FP32 4096^3 library matrix products (PMC MFMA-counter "
                         "calibration: 3 x 137.4 GFLOP each)")
    ap.add_argument("--workload", choices=["hlu", "batch", "serve", "build", "trace", "ops"], default="hlu",
                    help="hlu: the N=16384 factor+solve step (default); batch: many matrices in one launch "
                         "with autotuned leaves; serve: Poisson request stream; build: H-matrix construction "
                         "on the device")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--host-build", action="store_true", help="serve: build the matrices on the host")
    ap.add_argument("--serve-policy", choices=["sjf", "fifo"], default="sjf")
    ap.add_argument("--no-batch", action="store_true", help="hlu: skip the batched-configuration record")
    ap.add_argument("--no-serve", action="store_true", help="hlu: skip the serving-configuration record")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8}))
        return
    if args.calibrate_mfma:
        out = {}
        for dt in (torch.float64, torch.float32):
            a = torch.ones(4096, 4096, dtype=dt, device="cuda")
            b = torch.ones(4096, 4096, dtype=dt, device="cuda")
            t0 = time.perf_counter()
            for _ in range(3):
                c = a @ b
            torch.cuda.synchronize()
            out[str(dt)] = {"flops": 3 * 2 * 4096 ** 3, "seconds": time.perf_counter() - t0, "c00": float(c[0, 0])}
        print(json.dumps(out))
        return
    pp = importlib.import_module("pastel-palettes_amd")
    ctx = pp.Context(0)
    workers = args.workers or min(ctx.cu_count, 256)
    if args.workload != "hlu":
        wl = importlib.import_module("pastel-palettes_amd.workloads")
        if args.workload == "batch":
            out = wl.run_batch(pp, torch, ctx, count=args.batch_count, steps=args.steps, warmup=args.warmup,
                               log=log)
        elif args.workload == "trace":
            out = wl.run_trace(pp, torch, ctx, n=args.n, levels=args.levels, rank=args.rank, log=log)
        elif args.workload == "ops":
            out = wl.run_ops(pp, torch, ctx, log=log)
        elif args.workload == "build":
            out = wl.run_build(pp, torch, ctx, n=args.n, levels=args.levels, rank=args.rank, steps=args.steps,
                               warmup=args.warmup, log=log)
        else:
            out = wl.run_serve(pp, torch, ctx, duration=args.serve_seconds, slots=args.serve_slots,
                               device_build=not args.host_build, policy=args.serve_policy, log=log)
        print(json.dumps(out), flush=True)
        return

    t0 = time.perf_counter()
    hm, kern, arena = pp.build_config(1, args.n, levels=args.levels, rank=args.rank)
    b = pp.rhs(1, args.n, args.nrhs)
    log(f"built H-matrix n={args.n} storage={arena.size} in {time.perf_counter() - t0:.1f}s; workers={workers}")
    arena_dev = torch.from_numpy(arena).cuda()

    res = {}
    xs = {}
    precs = {"fp64": pp.FP64, "mixed": pp.MIXED}
    for name in (["fp64", "mixed"] if args.only == "both" else [args.only]):
        log(f"[bench] {name} steps={args.steps} warmup={args.warmup}")
        res[name], xs[name] = run_precision(pp, torch, ctx, hm, arena_dev, b, precs[name], args.steps, args.warmup,
                                            workers, args.wavefronts, args.nrhs, args.tol)
        log(f"[bench] {name}: {res[name]}")

    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    value = args.steps / r["seconds"]
    prog = pp.Program.factor(hm, workers)
    leaf, upd = factor_flop_split(prog)
    if head == "mixed":
        peak = (leaf + upd) / (leaf / PEAK_TFLOPS["fp32"] + upd / PEAK_TFLOPS["bf16"])
        dtype = "bf16/f32 mfma + f64 refinement"
    else:
        peak = PEAK_TFLOPS["fp64"]
        dtype = "f64"
    achieved = r["factor_flops"] / (r["factor_kernel_ms"] * 1e-3) / 1e12
    t_lb = lower_bound_ms(prog, head == "mixed")

    cpu = cpu0 = None
    errs = None
    if not args.no_cpu:
        import oracle
        threads = min(len(os.sched_getaffinity(0)), 16)
        log(f"[bench] cpu baseline on {threads} threads ({cpu_model()}, nproc {os.cpu_count()})")
        sec, reps, cres, cit, xc = cpu_baseline(pp, oracle, hm, arena, b, threads, args.tol)
        cpu = {"value": 1.0 / sec, "unit": "factor+solve/s", "cores": threads, "kind": "port",
               "sample": f"{reps} x (FP64 factorization + {args.nrhs}-rhs solve + FP64 refinement to "
                         f"{args.tol:g}: {cit} iterations, residual {cres:.2e}) of the same n={args.n} "
                         f"HODLR streams by the oracle executor on {threads} threads; host {cpu_model()}, "
                         f"nproc {os.cpu_count()}"}
        # every GPU solution against the FP64 oracle's (refined) solution
        errs = {name: float(np.linalg.norm(xs[name] - xc) / np.linalg.norm(xc)) for name in xs}
        # the reference's CPU configuration: n = 1024, 1-D exp kernel, all
        # blocks dense (leaf 128), one right-hand side
        hm0, _, arena0 = pp.build_config(0)
        b0 = pp.rhs(0, hm0.n, 1)
        sec0, reps0, cres0, cit0, _ = cpu_baseline(pp, oracle, hm0, arena0, b0, threads, args.tol, min_seconds=3.0,
                                                   max_reps=50)
        cpu0 = {"value": 1.0 / sec0, "unit": "factor+solve/s", "cores": threads, "kind": "port",
                "flops": int(pp.dense_lu_flops(hm0.n)),
                "gflops_factor_flops_per_step_time": pp.dense_lu_flops(hm0.n) / sec0 / 1e9,
                "sample": f"{reps0} x (FP64 factorization + 1-rhs solve + FP64 refinement to {args.tol:g}: "
                          f"{cit0} iterations, residual {cres0:.2e}) of n={hm0.n}, 1-D exp kernel, all-dense "
                          f"leaf-128 streams by the oracle executor on {threads} threads"}

    line = {
        "metric": "H-LU factor+solve/sec & p99 latency at residual <=1e-10, mixed vs plain FP64",
        "value": value,
        "unit": "factor+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["per_step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded points, kernel matrix compressed to HODLR; seeded N(0,1) rhs)",
        "config": {"workload": f"hodlr n={args.n} levels={args.levels} leaf={args.n >> args.levels} "
                               f"rank={args.rank} nrhs={args.nrhs} tol={args.tol}",
                   "workers": workers, "wavefronts_per_inst": args.wavefronts},
        "p50_ms": r["p50_ms"],
        "p99_ms": r["p99_ms"],
        "residual": r["residual"],
        "refinement_iterations": r["iterations"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(
                         f"vm_factor_kernel_{head}_hodlr_n{args.n}_l{args.levels}_r{args.rank}"),
                     "kernel": "vm_factor_kernel (factorization launch)", "factor_kernel_ms": r["factor_kernel_ms"],
                     "factor_flops": r["factor_flops"],
                     "t_lb_ms": t_lb, "t_lb_frac": t_lb / r["factor_kernel_ms"],
                     "t_lb_note": "sum over the H-LU instructions of max(FLOPs / MFMA peak, algorithmic bytes / "
                                  "8 TB/s); the inverse construction in the same launch is not counted"},
        "cpu_baseline": cpu,
        "cpu_baseline_n1024_dense": cpu0,
        "error_vs_fp64_oracle": errs,
        "summary": summary_columns(pp, hm, prog, r["factor_flops"], r["factor_kernel_ms"]),
    }
    traffic = line["roofline"]["traffic"]
    line["roofline"]["hbm_gbps"] = traffic / (r["factor_kernel_ms"] * 1e-3) / 1e9 if traffic else None
    line["roofline"]["mfma_busy_frac"] = measured_mfma_busy(head)
    line["roofline"]["counters_note"] = ("traffic and mfma_busy_frac are read from the committed rocprofv3 PMC "
                                         "summaries of vm_factor_kernel (profiles/); hbm_gbps = traffic / the "
                                         "factorization launch time measured here")
    if "fp64" in res:
        f = res["fp64"]
        line["fp64"] = {"value": args.steps / f["seconds"], "ms_per_step": f["per_step_ms"],
                        "factor_kernel_ms": f["factor_kernel_ms"], "residual": f["residual"],
                        "refinement_iterations": f["iterations"], "p50_ms": f["p50_ms"], "p99_ms": f["p99_ms"],
                        "achieved_tflops": f["factor_flops"] / (f["factor_kernel_ms"] * 1e-3) / 1e12,
                        "frac_of_fp64_peak": f["factor_flops"] / (f["factor_kernel_ms"] * 1e-3) / 1e12 /
                        PEAK_TFLOPS["fp64"]}
    if "fp64" in res and "mixed" in res:
        line["mixed_over_fp64"] = line["value"] / line["fp64"]["value"]
        line["mixed_vs_fp64_solution_rel_diff"] = float(np.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                                        np.linalg.norm(xs["fp64"]))
    # the batched and serving configurations, each a bounded run of its own
    # workload (bench.py --workload batch|serve runs them alone)
    wl = importlib.import_module("pastel-palettes_amd.workloads")
    if not args.no_batch:
        log("[bench] batch record")
        bt = wl.run_batch(pp, torch, ctx, count=args.batch_count, steps=2, warmup=1, log=log)
        line["batch"] = {k: bt[k] for k in ("workload", "matrices_per_s", "ms_per_batch", "factor_ms", "factor_tflops",
                                            "residual", "iterations", "converged", "leaf_choices", "solve_mode",
                                            "tuned_buckets")}
    if not args.no_serve:
        log("[bench] serve record")
        sv = wl.run_serve(pp, torch, ctx, duration=args.serve_seconds, slots=args.serve_slots,
                          policy=args.serve_policy, log=log)
        line["serve"] = {k: v for k, v in sv.items() if k not in ("sizes",)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
