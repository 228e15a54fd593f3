"""Per-instruction trace of a factorization launch, read back end to end
(reference: clock64() stamps per instruction, src/kernel.cu:1389-1390, read
by instructions_scheduler::analyzeClocks, rts/scheduler.cpp:494-544):
one stamp per stream entry of every worker plus one at its end, in order on
each worker; the reduction to per-operation totals and to the critical
path (the chain of instructions that set the launch's end) is consistent
with the stamps."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_factor_trace_readout(pp, ctx):
    from importlib import import_module
    tr = import_module("pastel-palettes_amd.trace")
    n, levels, rank, workers = 2048, 3, 16, 16
    hm, kern, arena = pp.build_config(1, n, levels=levels, rank=rank)
    solver = pp.Solver(ctx, hm, arena, precision=pp.FP64, workers=workers, nrhs_max=1)
    try:
        fs = pp.lib().pp_solver_factor_stream(solver.h)
        solver.factor()
        pp.check(pp.lib().pp_stream_enable_trace(ctx.h, fs, 1))
        solver.factor()
        n_vals = pp.check(pp.lib().pp_read_trace(ctx.h, fs, None, 0))
        stamps = np.zeros(n_vals, np.uint64)
        pp.lib().pp_read_trace(ctx.h, fs, stamps.ctypes.data_as(C.POINTER(C.c_uint64)), n_vals)
    finally:
        solver.close()
    prog = pp.Program.factor_inv_batch([hm], workers)
    streams = prog.streams()
    # one stamp per entry of every worker plus one at its end, non-decreasing
    # on each worker (the finish entry ends a stream)
    pos = 0
    for s in streams:
        ents = tr._entries(s)
        k = len(ents) + 1
        seg = stamps[pos:pos + k].astype(np.int64)
        done = len(ents) - 1 if ents and ents[-1][0] == "finish" else len(ents)
        assert np.all(seg[:done + 1] > 0), "unstamped entry"
        assert np.all(np.diff(seg[:done + 1]) >= 0), "stamps out of order on a worker"
        pos += k
    assert pos == n_vals
    summ = tr.summary(prog, stamps)
    leaves = sum(1 for lf in hm.leaves() if lf["type"] == "dense")
    assert summ["ops"]["getrf"]["count"] == leaves
    assert summ["span_us"] > 0 and 0.0 < summ["busy_fraction"] <= 1.0
    path, by_op = tr.critical_path(prog, stamps)
    assert path and path[0][1] == 1  # the chain starts at the first leaf's getrf
    # every entry of the path starts after its predecessor ended, and the
    # last one ends the launch
    for a, b in zip(path, path[1:]):
        assert b[2] >= a[3] - 1e-6
    assert abs(path[-1][3] - summ["span_us"]) < 1e-6
    assert sum(by_op.values()) <= summ["span_us"] + 1e-6
    assert by_op.get("getrf", 0.0) > 0.0
