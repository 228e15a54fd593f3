"""Per-instruction timing from the VM's entry timestamps.

The VM records a 100 MHz wall-clock stamp at every stream entry of every
worker (reference: clock64() per instruction, src/kernel.cu:1389-1390, shown
by instructions_scheduler::analyzeClocks, rts/scheduler.cpp:494-544).  This
module turns one traced launch into per-instruction (start, end) intervals,
per-operation totals, worker utilisation and a text Gantt chart.
"""
import collections

import numpy as np

OP_NAMES = {0: "nop", 1: "getrf", 2: "trsml", 3: "trsmr", 4: "gemm", 5: "gemm_plus", 6: "gemm_3x", 7: "gemm_4x",
            8: "accum", 9: "trsmu", 10: "trmm"}
TICK_US = 0.01  # 100 MHz


def _entries(stream):
    out, i = [], 0
    lens = {0: 3, 1: 8, 2: 13, 3: 13, 4: 17, 5: 17, 6: 23, 7: 29, 8: 21, 9: 13, 10: 11}
    while i < len(stream):
        op = int(stream[i])
        if op == 0:
            out.append(("finish", -1, -1))
            break
        if op == 2:
            out.append(("wait", int(stream[i + 1]), -1))
            i += 3
        else:
            o = int(stream[i + 2])
            out.append(("exec", int(stream[i + 1]), o))
            i += lens[o]
    return out


def intervals(program, trace):
    """List of per-entry records (worker, kind, inst, op, t0_us, t1_us)."""
    recs = []
    pos = 0
    t_min = None
    for w, stream in enumerate(program.streams()):
        ents = _entries(stream)
        n = len(ents)
        stamps = trace[pos:pos + n + 1].astype(np.int64)
        pos += n + 1
        for k, (kind, inst, op) in enumerate(ents):
            if kind == "finish":
                break
            t0, t1 = stamps[k], stamps[k + 1]
            recs.append((w, kind, inst, op, t0, t1))
            t_min = t0 if t_min is None else min(t_min, t0)
    return [(w, k, i, o, (a - t_min) * TICK_US, (b - t_min) * TICK_US) for w, k, i, o, a, b in recs]


def summary(program, trace):
    recs = intervals(program, trace)
    by_op = collections.defaultdict(lambda: [0, 0.0])
    busy = collections.defaultdict(float)
    wait = collections.defaultdict(float)
    end = 0.0
    for w, kind, inst, op, t0, t1 in recs:
        end = max(end, t1)
        if kind == "exec":
            by_op[OP_NAMES.get(op, str(op))][0] += 1
            by_op[OP_NAMES.get(op, str(op))][1] += t1 - t0
            busy[w] += t1 - t0
        else:
            wait[w] += t1 - t0
    W = len(program.streams())
    return {
        "span_us": end,
        "ops": {k: {"count": v[0], "total_us": v[1], "mean_us": v[1] / max(v[0], 1)} for k, v in by_op.items()},
        "busy_fraction": sum(busy.values()) / (W * end) if end > 0 else 0.0,
        "wait_us_total": sum(wait.values()),
    }


def gantt(program, trace, width=100, workers=None):
    """Text chart: '#' executing, '.' waiting, ' ' idle (one row per worker)."""
    recs = intervals(program, trace)
    end = max(r[5] for r in recs) if recs else 1.0
    W = len(program.streams()) if workers is None else workers
    rows = [[" "] * width for _ in range(W)]
    for w, kind, inst, op, t0, t1 in recs:
        if w >= W:
            continue
        a, b = int(t0 / end * (width - 1)), int(t1 / end * (width - 1))
        for c in range(a, b + 1):
            if kind == "exec" or rows[w][c] == " ":
                rows[w][c] = "#" if kind == "exec" else "."
    return "\n".join(f"{w:4d} |{''.join(r)}|" for w, r in enumerate(rows))


def critical_path(program, trace):
    """Chain of instructions that determined the traced launch's end time.

    Walks back from the last instruction to finish: at each step the
    predecessor (DAG dependency, or the previous instruction of the same
    worker) that finished last is the one the instruction waited for.
    Returns (path, by_op) where path is a list of (inst, op, t0_us, t1_us)
    from first to last and by_op sums the path's execution time per
    operation; the gaps between consecutive path entries are signalling and
    scheduling latency."""
    recs = [r for r in intervals(program, trace) if r[1] == "exec"]
    if not recs:
        return [], {}
    info = {r[2]: r for r in recs}  # inst -> (w, kind, inst, op, t0, t1)
    prev_on_worker = {}
    last = {}
    for w, kind, inst, op, t0, t1 in sorted(recs, key=lambda r: (r[0], r[4])):
        if w in last:
            prev_on_worker[inst] = last[w]
        last[w] = inst
    cur = max(recs, key=lambda r: r[5])[2]
    path = []
    while cur is not None:
        w, kind, inst, op, t0, t1 = info[cur]
        path.append((inst, op, t0, t1))
        cands = [p for p, _ in program.preds(inst) if p in info]
        if inst in prev_on_worker:
            cands.append(prev_on_worker[inst])
        cur = max(cands, key=lambda p: info[p][5]) if cands else None
    path.reverse()
    by_op = collections.defaultdict(float)
    for inst, op, t0, t1 in path:
        by_op[OP_NAMES.get(op, str(op))] += t1 - t0
    return path, dict(by_op)


def schedule_summary(program):
    """The reference scheduler's summary columns for a program's streams
    (instructions_scheduler, rts/scheduler.cpp:406-427): instruction count,
    the heaviest worker's share of the FLOPs, utilisation (total FLOPs over
    workers x heaviest worker's FLOPs) and cross-worker waits per
    instruction."""
    info = program.info()
    per_worker = []
    for s in program.streams():
        per_worker.append(sum(program.inst_flops(i) for kind, i, _ in _entries(s) if kind == "exec"))
    total = float(sum(per_worker))
    heaviest = float(max(per_worker)) if per_worker else 0.0
    n = info["n_insts"]
    return {
        "instructions": n,
        "workers": info["workers"],
        "heaviest_worker_pct": 100.0 * heaviest / total if total else 0.0,
        "utilization_pct": 100.0 * total / heaviest / len(per_worker) if heaviest else 0.0,
        "comm_per_inst": info["comm_waits"] / n if n else 0.0,
    }
