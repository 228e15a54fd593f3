"""Batched and serving benchmark drivers (the many-matrix workloads beside the
single N = 16384 H-LU step of bench.py).

batch: `count` independent HODLR matrices, N log-uniform in [512, 8192]
    rounded to a multiple of 256, rank drawn from {16, 24, 32}, 4 right-hand
    sides each, leaf order per size chosen by the autotuner; one batched
    solver (one persistent factor launch replayed from a captured hipGraph,
    one solve launch per refinement sweep), FP32 factorization with FP64
    refinement to the tolerance.
serve: a Poisson stream of factor + solve requests over a pool of resident
    matrices with N log-uniform in [1024, 32768], 1-64 right-hand sides, 30 %
    of requests reusing a resident factorization; mixed precision with FP64
    refinement; HBM-budgeted admission.  The open-loop rate is a fraction of
    the peak rate measured closed-loop on the same pool.

Batch matrices are built on the host (not timed) by a pool of spawned
processes; the serving pool is built on the device (pp_hmatrix_build_kernel)
unless the host builder is asked for.
"""
import math
import multiprocessing as mp
import os
import time

import numpy as np


def _build_arena(job):
    """(n, levels, rank, point seed) -> FP64 arena; runs in a worker process."""
    import importlib
    import sys
    n, levels, rank, seed = job
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    pp = importlib.import_module("pastel-palettes_amd")
    hm = pp.HMatrix("hodlr", n, levels=levels, rank=rank)
    return pp.fill_arena(hm, pp.Kernel(pp.points_2d_morton(seed, n), "inv2d", n / 4.0))


def build_arenas(jobs, procs):
    if procs <= 1 or len(jobs) == 1:
        return [_build_arena(j) for j in jobs]
    # one BLAS thread per worker process: spinning BLAS thread pools in every
    # worker oversubscribe the cores many times over
    keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {k: os.environ.get(k) for k in keys}
    for k in keys:
        os.environ[k] = "1"
    try:
        with mp.get_context("spawn").Pool(procs) as pool:
            # largest first, so the long builds do not trail at the end
            order = sorted(range(len(jobs)), key=lambda i: -jobs[i][0])
            res = pool.map(_build_arena, [jobs[i] for i in order], chunksize=1)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    out = [None] * len(jobs)
    for i, a in zip(order, res):
        out[i] = a
    return out


def log_uniform_sizes(rng, count, lo, hi, multiple):
    s = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    return [int(min(hi, max(lo, multiple * round(v / multiple)))) for v in s]


def host_procs():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def run_batch(pp, torch, ctx, count=512, nrhs=4, steps=3, warmup=1, tol=1e-10, precision=None, log=print):
    from .tune import Autotuner
    precision = pp.FP32 if precision is None else precision
    rng = np.random.default_rng(3000)
    sizes = log_uniform_sizes(rng, count, 512, 8192, 256)
    ranks = [int(r) for r in rng.choice([16, 24, 32], count)]
    workers = min(ctx.cu_count, 256)
    t0 = time.perf_counter()
    tuner = Autotuner(pp, torch, ctx, precision, rank=32, nrhs=nrhs, tol=tol, workers=workers, log=log)
    choices = [tuner.choose(n) for n in sizes]
    tune_s = time.perf_counter() - t0
    jobs = [(n, c["levels"], r, 100 + i) for i, (n, c, r) in enumerate(zip(sizes, choices, ranks))]
    log(f"[batch] building {count} matrices on {host_procs()} host processes")
    t0 = time.perf_counter()
    arenas = build_arenas(jobs, host_procs())
    build_s = time.perf_counter() - t0
    log(f"[batch] built in {build_s:.1f} s; creating the batched solver")
    hms = [pp.HMatrix("hodlr", n, levels=lv, rank=r) for n, lv, r, _ in jobs]
    t0 = time.perf_counter()
    from .tune import choose_solve_mode
    # the mode is timed on the whole batch: a sample of it is less throughput
    # bound than the batch and favours the mode with the cheapest solves
    mode, mode_cost = choose_solve_mode(pp, torch, ctx, hms, arenas, precision, nrhs=nrhs, tol=tol, workers=workers,
                                        sample=len(hms), reps=1, log=log)
    tune_s += time.perf_counter() - t0
    log(f"[batch] solve mode {mode}")
    t0 = time.perf_counter()
    solver = pp.Solver.batch(ctx, hms, arenas, precision=precision, workers=workers, nrhs_max=nrhs, solve_mode=mode)
    solver.use_graph(True)
    setup_s = time.perf_counter() - t0
    log(f"[batch] solver ready in {setup_s:.1f} s ({int(sum(sizes))} rows)")
    del arenas
    b = np.concatenate([pp.rhs(4 + (i % 64), n, nrhs) for i, n in enumerate(sizes)], axis=0)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    stream = torch.cuda.Stream()
    for _ in range(warmup):
        solver.factor(stream.cuda_stream)
        solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=tol, max_iter=60, hip_stream=stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fms, reps = [], []
    for _ in range(steps):
        solver.factor(stream.cuda_stream)
        fms.append(solver.last_factor_ms())
        reps.append(solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=tol, max_iter=60,
                                 hip_stream=stream.cuda_stream))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flops = solver.factor_flops
    leaf_hist = {}
    for c in choices:
        leaf_hist[c["leaf"]] = leaf_hist.get(c["leaf"], 0) + 1
    out = {
        "workload": f"batch of {count} HODLR matrices, N log-uniform [512, 8192] (multiple of 256), "
                    f"rank 16/24/32, {nrhs} rhs, FP32 factor + FP64 refinement to {tol:g}, hipGraph factor launch",
        "matrices_per_s": count * steps / dt,
        "ms_per_batch": 1e3 * dt / steps,
        "factor_ms": float(np.mean(fms)),
        "factor_tflops": flops / (float(np.mean(fms)) * 1e-3) / 1e12,
        "total_rows": int(sum(sizes)),
        "residual": max(r["residual"] for r in reps),
        "iterations": max(r["iterations"] for r in reps),
        "converged": all(r["converged"] for r in reps),
        "leaf_choices": leaf_hist,
        "solve_mode": {0: "hierarchical inverse", 1: "leaf inverses", 2: "substitution"}[mode],
        "solve_mode_ms_per_matrix": {str(k): v for k, v in mode_cost.items()},
        "tuned_buckets": {k: {"leaf": v["leaf"], "ms_per_matrix": v["ms_per_matrix"],
                              "candidates": v["candidates"]} for k, v in tuner.cache.items()},
        "tune_s": tune_s, "build_s": build_s, "setup_s": setup_s, "workers": workers,
    }
    solver.close()
    return out


def _size_bucket(n):
    return 1 << int(round(math.log2(n)))


def run_serve(pp, torch, ctx, keys=32, duration=20.0, load=0.8, burst=64, slots=4, tol=1e-10, budget_gb=200,
              device_build=True, policy="sjf", share=True, log=print):
    """Serving benchmark: `keys` matrices (N log-uniform in [1024, 32768],
    the largest set to 32768) built on the device; every matrix factored and
    solved once alone (its alone service times, factor + solve and solve
    alone); the closed-loop peak over a burst of the request mix; then a
    Poisson stream at `load` times that peak for `duration` seconds.
    Reports overall p50 / p99 latency, per-size-bucket latencies, and the
    slowdown (latency over the request's alone service time), whose spread
    is what the admission policy controls: the spread of the raw latencies is
    bounded below by the mix's service times, which grow with N (a chain of
    diagonal-leaf factorizations)."""
    from . import serve
    rng = np.random.default_rng(5000)
    sizes = log_uniform_sizes(rng, keys, 1024, 32768, 256)
    sizes[int(np.argmax(sizes))] = 32768
    jobs = [(n, max(1, int(round(math.log2(n / 256)))), 32, 200 + i) for i, n in enumerate(sizes)]
    log(f"[serve] building {keys} matrices (N = {sizes}) on the {'device' if device_build else 'host'}")
    t0 = time.perf_counter()
    matrices = {}
    if device_build:
        # construction on the device: points uploaded, arena filled in place
        for i, (n, lv, r, seed) in enumerate(jobs):
            hm = pp.HMatrix("hodlr", n, levels=lv, rank=r)
            pts = torch.from_numpy(np.ascontiguousarray(pp.points_2d_morton(seed, n))).cuda()
            arena = torch.empty(hm.storage, dtype=torch.float64, device="cuda")
            hm.build_on_device(ctx, pp.KERNEL_INV2D, pts.data_ptr(), 1e-3, n / 4.0, arena.data_ptr())
            matrices[f"m{i}"] = (hm, arena)
    else:
        arenas = build_arenas(jobs, host_procs())
        for i, ((n, lv, r, _), a) in enumerate(zip(jobs, arenas)):
            matrices[f"m{i}"] = (pp.HMatrix("hodlr", n, levels=lv, rank=r), torch.from_numpy(a).cuda())
        del arenas
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"[serve] built in {build_s:.1f} s")
    names = sorted(matrices)
    workers = min(ctx.cu_count, 256) // slots
    server = serve.Server(pp, torch, ctx, matrices, pp.MIXED, int(budget_gb * 1e9), slots=slots,
                          workers_per_slot=workers, tol=tol, policy=policy,
                          share=serve.default_share if share else None)

    def make_rhs(key, nrhs, k):
        return pp.rhs(7 + k % 97, matrices[key][0].n, nrhs)

    try:
        # every factorization made resident (solver construction included),
        # then each matrix's alone service times: factor + solve, and solve
        # with the resident factors, one request on the device at a time
        t0 = time.perf_counter()
        serve.replay(server, [(0.0, k, 1, False) for k in names], make_rhs, realtime=False)
        alone = {}
        for k in names:
            for reuse in (False, True):
                r, _ = serve.replay(server, [(0.0, k, 1, reuse)], make_rhs, realtime=False)
                alone[(k, reuse)] = r[0].done - r[0].start
        setup_s = time.perf_counter() - t0
        log(f"[serve] all matrices resident ({setup_s:.1f} s)")
        # peak: a closed burst (all requests queued at once) of the same mix
        tr = serve.poisson_trace(1e9, burst / 1e9, names, seed=11)[:burst]
        reqs, wall = serve.replay(server, [(0.0, k, nr, ru) for _, k, nr, ru in tr], make_rhs, realtime=False)
        peak = serve.summarize(reqs, wall)
        rate = load * peak["throughput_rps"]
        log(f"[serve] peak {peak['throughput_rps']:.1f} req/s; open loop at {rate:.1f} req/s for {duration:.0f} s")
        trace = serve.poisson_trace(rate, duration, names, seed=12)
        reqs, wall = serve.replay(server, trace, make_rhs, realtime=True)
        s = serve.summarize(reqs, wall)
        ok = [q for q in reqs if q.error is None]
        by_size = {}
        for q in ok:
            by_size.setdefault(_size_bucket(matrices[q.key][0].n), []).append(q.latency)
        per_n = {str(n): {"count": len(v), "p50_ms": 1e3 * float(np.percentile(v, 50)),
                          "p99_ms": 1e3 * float(np.percentile(v, 99))} for n, v in sorted(by_size.items())}
        slow = np.array([q.latency / alone[(q.key, not q.factored)] for q in ok])
        svc = np.array([alone[(q.key, not q.factored)] for q in ok])
    finally:
        server.close()
    return {
        "workload": f"serving: Poisson requests over {keys} resident HODLR matrices, N log-uniform [1024, 32768], "
                    f"rank 32, leaf ~256, 1-64 rhs, 30% reuse, mixed precision + FP64 refinement to {tol:g}, "
                    f"{slots} streams sharing {slots * workers} workers ({'worker share by N' if share else 'fixed slots'}), "
                    f"{policy} admission, HBM budget {budget_gb} GB, matrices built on the "
                    f"{'device' if device_build else 'host'}",
        "sizes": sizes,
        "peak_rps": peak["throughput_rps"],
        "rate_rps": rate,
        "throughput_rps": s["throughput_rps"],
        "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
        "p99_over_p50": s["p99_ms"] / s["p50_ms"] if s["p50_ms"] else None,
        "alone_service_p50_ms": 1e3 * float(np.percentile(svc, 50)),
        "alone_service_p99_ms": 1e3 * float(np.percentile(svc, 99)),
        "alone_service_p99_over_p50": float(np.percentile(svc, 99) / np.percentile(svc, 50)),
        "slowdown_p50": float(np.percentile(slow, 50)), "slowdown_p99": float(np.percentile(slow, 99)),
        "slowdown_p99_over_p50": float(np.percentile(slow, 99) / np.percentile(slow, 50)),
        "requests": s["requests"], "errors": s["errors"], "factored": s["factored"],
        "max_residual": s["max_residual"], "converged": s["converged"],
        "per_size": per_n, "evictions": server.evictions, "build_s": build_s, "setup_s": setup_s,
    }


def _hmatrix_rows(hm, arena, rows, n):
    """Rows `rows` of the matrix an arena represents (dense n-vectors)."""
    out = np.zeros((len(rows), n))
    pos = {r: i for i, r in enumerate(rows)}
    for lf in hm.leaves():
        r0, c0, ny, nx = lf["row0"], lf["col0"], lf["ny"], lf["nx"]
        hit = [r for r in rows if r0 <= r < r0 + ny]
        for r in hit:
            t = r - r0
            if lf["type"] == "dense":
                out[pos[r], c0:c0 + nx] = arena[lf["off"] + t * nx:lf["off"] + (t + 1) * nx]
            else:
                k = lf["rank"]
                u = arena[lf["off"] + t * k:lf["off"] + (t + 1) * k]
                v = arena[lf["off_v"]:lf["off_v"] + nx * k].reshape(nx, k)
                out[pos[r], c0:c0 + nx] = v @ u
    return out


def run_build(pp, torch, ctx, n=16384, levels=6, rank=32, steps=5, warmup=1, sample_rows=64, log=print):
    """H-matrix construction on the device (2-D 1/(r + 1e-3) kernel of the
    benchmark configuration) against the host builder (randomised SVD with
    multithreaded BLAS); compression error on sampled rows of the kernel."""
    hm = pp.HMatrix("hodlr", n, levels=levels, rank=rank)
    kern = pp.config_kernel(1, n)
    pts = torch.from_numpy(np.ascontiguousarray(kern.pts, np.float64)).cuda()
    arena = torch.zeros(hm.storage, dtype=torch.float64, device="cuda")

    def build():
        hm.build_on_device(ctx, pp.KERNEL_INV2D, pts.data_ptr(), 1e-3, kern.shift, arena.data_ptr())

    for _ in range(warmup):
        build()
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        build()
        times.append(1e3 * (time.perf_counter() - t0))
    times.sort()
    log(f"[build] device: {times}")
    t0 = time.perf_counter()
    host = pp.fill_arena(hm, kern)
    host_s = time.perf_counter() - t0
    log(f"[build] host randomised SVD: {host_s:.2f} s")
    dev = arena.cpu().numpy()
    rng = np.random.default_rng(11)
    rows = sorted(rng.choice(n, sample_rows, replace=False).tolist())
    exact = np.stack([kern.block(r, 1, 0, n)[0] for r in rows])
    err = lambda a: float(np.linalg.norm(exact - _hmatrix_rows(hm, a, rows, n)) / np.linalg.norm(exact))
    stored = sum(l["ny"] * l["nx"] if l["type"] == "dense" else (l["ny"] + l["nx"]) * l["rank"] for l in hm.leaves())
    med = times[len(times) // 2]
    return {
        "workload": f"device H-matrix construction: hodlr n={n} levels={levels} rank={rank}, 2-D 1/(r+1e-3) "
                    f"kernel (Morton points, shift n/4); dense leaves evaluated, low-rank leaves by cross "
                    f"approximation of rank max(4r,64) truncated to r",
        "ms_per_build": med, "builds_per_s": 1e3 / med, "ms_min": times[0], "ms_max": times[-1],
        "stored_entries": int(stored), "stored_gb_per_s": stored * 8 / (med * 1e-3) / 1e9,
        "host_build_s": host_s, "host_builder": "numpy randomised SVD (oversampling 10, one power iteration), "
                                                "multithreaded BLAS",
        "speedup_vs_host": host_s * 1e3 / med,
        "sampled_rows": sample_rows,
        "compression_error_device": err(dev), "compression_error_host": err(host),
    }


def run_trace(pp, torch, ctx, n=16384, levels=6, rank=32, precision=None, workers=None, log=print):
    """One traced factorization launch of the benchmark matrix (the solver's
    factor program: H-LU plus the inverse construction) per precision:
    per-entry wall-clock stamps of every worker (reference: clock64() per
    instruction, src/kernel.cu:1389-1390, read by analyzeClocks,
    rts/scheduler.cpp:494-544), reduced to the critical path -- the chain of
    instructions that set the launch's end -- with per-operation time on it
    and the signalling gaps between its entries."""
    from . import trace as tr
    precs = {"fp64": pp.FP64, "mixed": pp.MIXED} if precision is None else {precision: pp.PRECISIONS[precision]}
    hm, kern, arena = pp.build_config(1, n, levels=levels, rank=rank)
    workers = workers or min(ctx.cu_count, 256)
    out = {"workload": f"traced factorization launch, hodlr n={n} levels={levels} rank={rank}, {workers} workers"}
    for name, prec in precs.items():
        solver = pp.Solver(ctx, hm, arena, precision=prec, workers=workers, nrhs_max=1)
        fs = pp.lib().pp_solver_factor_stream(solver.h)
        solver.factor()  # warm
        pp.check(pp.lib().pp_stream_enable_trace(ctx.h, fs, 1))
        solver.factor()
        ms = solver.factor_stream_ms()
        n_vals = pp.check(pp.lib().pp_read_trace(ctx.h, fs, None, 0))
        stamps = np.zeros(n_vals, np.uint64)
        import ctypes as C
        pp.lib().pp_read_trace(ctx.h, fs, stamps.ctypes.data_as(C.POINTER(C.c_uint64)), n_vals)
        prog = _factor_program_of(pp, solver, hm, workers)
        summ = tr.summary(prog, stamps)
        path, by_op = tr.critical_path(prog, stamps)
        gaps = sum(max(0.0, b[2] - a[3]) for a, b in zip(path, path[1:]))
        counts = {}
        for inst, op, t0, t1 in path:
            k = tr.OP_NAMES.get(op, str(op))
            counts[k] = counts.get(k, 0) + 1
        out[name] = {
            "launch_ms": ms, "span_us": summ["span_us"], "busy_fraction": summ["busy_fraction"],
            "critical_path": {"entries": len(path), "exec_us": sum(by_op.values()), "gap_us": gaps,
                              "by_op": {k: {"count": counts[k], "total_us": v, "mean_us": v / counts[k]}
                                        for k, v in sorted(by_op.items(), key=lambda kv: -kv[1])}},
            "ops": summ["ops"],
        }
        solver.close()
    return out


def _factor_program_of(pp, solver, hm, workers):
    """The factor program a single-matrix solver launches (factorization
    followed by the hierarchical-inverse construction), rebuilt the same way
    so that its streams match the traced launch's."""
    return pp.Program.factor_inv_batch([hm], workers)


def run_ops(pp, torch, ctx, reps=50, log=print):
    """Latency of the operations on the benchmark's critical path, one
    workgroup each, executed `reps` times in place (pp_exec_single_timed):
    256 x 256 (and 128, 160) diagonal-leaf getrfs, the 256 x 32 trsm of a low-rank
    factor (left, and right with transposed storage), and the update of a
    leaf by a rank-32 product (gemm_4x in the ((A B) C) D and A ((B^T C) D^T)
    associations, and the 256 x 256 x 32 gemm_plus)."""
    rng = np.random.default_rng(1)
    res = {}
    for name, prec in (("fp64", pp.FP64), ("fp32", pp.FP32), ("mixed", pp.MIXED)):
        dt = torch.float64 if prec == pp.FP64 else torch.float32
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
        n, r = 256, 32
        a = rng.standard_normal((n, n)) + 2 * n * np.eye(n)
        lu = d(a)
        row = {}
        # getrf in place on a fresh copy per rep is not possible; the timed
        # reps refactor the factors, the same work at the same shape
        row["getrf_256"] = 1e3 * ctx.time_single(prec, [1, 0, 1, 0, 0, n, n, n], [lu.data_ptr()], reps=reps)
        for m in (128, 160):
            a_m = d(rng.standard_normal((m, m)) + 2 * m * np.eye(m))
            row[f"getrf_{m}"] = 1e3 * ctx.time_single(prec, [1, 0, 1, 0, 0, m, m, m], [a_m.data_ptr()], reps=reps)
        L = d(np.tril(rng.standard_normal((n, n)) / n, -1) + np.eye(n))
        B = d(rng.standard_normal((n, r)))
        row["trsml_256x32"] = 1e3 * ctx.time_single(prec, [1, 0, 2, 0, 1, 0, 0, r, n, n, r, n, 0],
                                                    [B.data_ptr(), L.data_ptr()], reps=reps)
        U = d(np.triu(rng.standard_normal((n, n)) / n) + 2 * np.eye(n))
        Bt = d(rng.standard_normal((n, r)))  # V factor: B^T stored (n x r)
        row["trsmr_t_256x32"] = 1e3 * ctx.time_single(prec, [1, 0, 3, 0, 1, 0, 0, n, r, n, r, n, 1],
                                                      [Bt.data_ptr(), U.data_ptr()], reps=reps)
        M = d(rng.standard_normal((n, n)))
        Ua, Vb = d(rng.standard_normal((n, r))), d(rng.standard_normal((n, r)))
        row["gemm_plus_256x256x32"] = 1e3 * ctx.time_single(
            prec, [1, 0, 5, 0, 1, 2, 0, 0, 0, n, n, r, n, r, r, 0, 1], [M.data_ptr(), Ua.data_ptr(), Vb.data_ptr()],
            reps=reps)
        C1, C2 = d(rng.standard_normal((r, r))), d(rng.standard_normal((n, r)))
        t1 = n * r
        row["gemm_4x_256x256_r32"] = 1e3 * ctx.time_single(
            prec, [1, 0, 7, 0, 1, 2, 3, 4, 0, 0, 0, 0, 0, n, n, r, r, r, n, r, r, r, r, 0, 0, 0, 1, 0,
                   (t1 + 31) // 32 * 32],
            [M.data_ptr(), Ua.data_ptr(), C1.data_ptr(), C1.data_ptr(), C2.data_ptr()], reps=reps)
        # the form on the factorization's critical path: M -= A ((B^T C) D^T),
        # B and C 256 x 32 (the inner product over the leaf's 256 rows)
        row["gemm_4x_256x256_r32_abc"] = 1e3 * ctx.time_single(
            prec, [1, 0, 7, 0, 1, 2, 3, 4, 0, 0, 0, 0, 0, n, n, r, n, r, n, r, r, r, r, 0, 1, 0, 1, 1, r * r],
            [M.data_ptr(), Ua.data_ptr(), Vb.data_ptr(), C2.data_ptr(), C2.data_ptr()], reps=reps)
        res[name] = row
        log(f"[ops] {name}: " + ", ".join(f"{k} {v:.1f} us" for k, v in row.items()))
    return {"workload": "latency of one critical-path operation on one workgroup (microseconds, mean of "
                        f"{reps} in-place executions)", **res}
