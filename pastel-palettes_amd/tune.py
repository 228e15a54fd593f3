"""Autotuner for batched H-LU launches: per matrix size, the leaf order of the
HODLR partition and the wavefronts per VM instruction.

A candidate is timed on the device the way a batch runs it: `copies` copies
of one matrix of the bucket's size are merged into a single batched solver
(one persistent factor launch, one solve launch), and the time of factor +
solve + FP64 refinement divided by `copies` is the candidate's cost per
matrix.  Sizes are bucketed by powers of two (in log scale); the choice for a
size is the cheapest candidate of its bucket.  Results can be kept in a JSON
cache file between runs.

Wavefronts per instruction: the VM kernel is instantiated for 4 wavefronts
(one 256-thread workgroup per worker) only, so that knob has one candidate
here; the leaf order is the one that moves the result.

The solve mode (hierarchical inverses formed in the factorization launch,
leaf inverses only, or substitution) is chosen for a whole batch by timing a
sample of its members in each mode.
"""
import json
import math
import os
import time

import numpy as np


def hodlr_levels(n, leaf):
    """Binary HODLR levels giving diagonal leaves closest to `leaf` rows."""
    return max(1, int(round(math.log2(max(n / float(leaf), 1.0)))))


class Autotuner:
    def __init__(self, pp, torch, ctx, precision, rank=32, nrhs=4, leaves=(128, 256), wavefronts=(4,), copies=8,
                 reps=3, tol=1e-10, workers=None, cache_path=None, log=None):
        self.pp, self.torch, self.ctx = pp, torch, ctx
        self.precision, self.rank, self.nrhs = precision, rank, nrhs
        self.leaves, self.wavefronts = tuple(leaves), tuple(wavefronts)
        self.copies, self.reps, self.tol = copies, reps, tol
        self.workers = workers or min(ctx.cu_count, 256)
        self.cache_path = cache_path
        self.log = log or (lambda *a: None)
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    @staticmethod
    def bucket(n):
        return 1 << max(9, int(round(math.log2(n))))

    def _key(self, b):
        return f"n{b}_r{self.rank}_p{self.precision}_k{self.nrhs}"

    def measure(self, n, leaf, wavefronts):
        """Milliseconds per matrix of factor + solve + refinement for a batch of
        `copies` matrices of order n with the given leaf target."""
        pp, torch = self.pp, self.torch
        levels = hodlr_levels(n, leaf)
        hm = pp.HMatrix("hodlr", n, levels=levels, rank=self.rank)
        # the candidate matrix is built on the device (the timing does not
        # depend on how its low-rank blocks were compressed)
        pts = torch.from_numpy(np.ascontiguousarray(pp.points_2d_morton(17, n))).cuda()
        dev = torch.empty(hm.storage, dtype=torch.float64, device="cuda")
        hm.build_on_device(self.ctx, pp.KERNEL_INV2D, pts.data_ptr(), 1e-3, n / 4.0, dev.data_ptr())
        s = pp.Solver.batch(self.ctx, [hm] * self.copies, [dev.data_ptr()] * self.copies, precision=self.precision,
                            workers=self.workers, wavefronts=wavefronts, nrhs_max=self.nrhs)
        b = np.concatenate([pp.rhs(3, n, self.nrhs)] * self.copies, axis=0)
        db = torch.from_numpy(b).cuda()
        dx = torch.zeros_like(db)
        try:
            s.factor()
            s.solve(db.data_ptr(), dx.data_ptr(), self.nrhs, tol=self.tol, max_iter=60)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(self.reps):
                s.factor()
                rep = s.solve(db.data_ptr(), dx.data_ptr(), self.nrhs, tol=self.tol, max_iter=60)
            torch.cuda.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / (self.reps * self.copies)
        finally:
            s.close()
        if not rep["converged"]:
            return float("inf"), levels
        return ms, levels

    def choose(self, n):
        """{'leaf', 'levels', 'wavefronts', 'ms_per_matrix', 'bucket', 'candidates'} for size n."""
        b = self.bucket(n)
        key = self._key(b)
        if key not in self.cache:
            cands = []
            for leaf in self.leaves:
                if b // leaf < 2:
                    continue
                for wf in self.wavefronts:
                    ms, levels = self.measure(b, leaf, wf)
                    self.log(f"[tune] n={b} leaf={leaf} levels={levels} wavefronts={wf}: {ms:.3f} ms/matrix")
                    cands.append({"leaf": leaf, "wavefronts": wf, "ms_per_matrix": ms})
            if not cands:
                cands.append({"leaf": min(self.leaves), "wavefronts": self.wavefronts[0], "ms_per_matrix": None})
            best = min(cands, key=lambda c: c["ms_per_matrix"] if c["ms_per_matrix"] is not None else 0.0)
            self.cache[key] = dict(best, bucket=b, candidates=cands)
            if self.cache_path:
                with open(self.cache_path, "w") as f:
                    json.dump(self.cache, f, indent=1)
        c = dict(self.cache[key])
        c["levels"] = hodlr_levels(n, c["leaf"])
        return c


def choose_solve_mode(pp, torch, ctx, hms, arenas, precision, nrhs=4, tol=1e-10, workers=None, sample=160,
                      modes=None, reps=2, log=None):
    """Solve mode with the least time per matrix for a sample (every k-th
    member) of a batch: returns (mode, {mode: ms per matrix})."""
    log = log or (lambda *a: None)
    workers = workers or min(ctx.cu_count, 256)
    modes = modes or (pp.SOLVE_HINV, pp.SOLVE_LEAFINV, pp.SOLVE_SUBST)
    step = max(1, len(hms) // sample)
    idx = list(range(0, len(hms), step))[:sample]
    sh, sa = [hms[i] for i in idx], [arenas[i] for i in idx]
    b = np.concatenate([pp.rhs(4 + (i % 64), sh[q].n, nrhs) for q, i in enumerate(idx)], axis=0)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    cost = {}
    for mode in modes:
        s = pp.Solver.batch(ctx, sh, sa, precision=precision, workers=workers, nrhs_max=nrhs, solve_mode=mode)
        try:
            s.factor()
            s.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=tol, max_iter=60)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                s.factor()
                rep = s.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=tol, max_iter=60)
            torch.cuda.synchronize()
            cost[mode] = 1e3 * (time.perf_counter() - t0) / (reps * len(sh)) if rep["converged"] else float("inf")
        finally:
            s.close()
        log(f"[tune] solve mode {mode}: {cost[mode]:.3f} ms/matrix over {len(sh)} sampled matrices")
    return min(cost, key=cost.get), cost
