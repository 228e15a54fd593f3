"""Host-side checkers for the full-size configuration tests (test helper):
the CPU oracle executing factor / solve instruction streams in FP64, the
oracle H-matrix-vector residual, and leaf-by-leaf factor comparison."""
import os

import numpy as np


def host_threads():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def leaf_products(hm, arena, base=0):
    """Dense leaves and low-rank products U V^T of the arena at `base`."""
    out = []
    for lf in hm.leaves():
        ny, nx = lf["ny"], lf["nx"]
        if lf["type"] == "dense":
            o = base + lf["off"]
            out.append(arena[o:o + ny * nx].reshape(ny, nx))
        else:
            r = lf["rank"]
            u = arena[base + lf["off"]:base + lf["off"] + ny * r].reshape(ny, r)
            v = arena[base + lf["off_v"]:base + lf["off_v"] + nx * r].reshape(nx, r)
            out.append(u @ v.T)
    return out


def worst_leaf_diff(hm, got, want, base_got=0):
    """Largest over leaves of max |got - want| / max |want| (leaf products)."""
    worst = 0.0
    for g, o in zip(leaf_products(hm, got, base_got), leaf_products(hm, want)):
        worst = max(worst, float(np.max(np.abs(g - o)) / max(np.max(np.abs(o)), 1e-300)))
    return worst


def oracle_factor(pp, orc, hm, arena, threads=None):
    """The FP64 factor arena the oracle computes from the factorization
    program's instructions (a schedule over `threads` workers: every
    instruction's arithmetic is independent of the schedule)."""
    threads = threads or host_threads()
    prog = pp.Program.factor(hm, threads)
    info = prog.info()
    fac = np.ascontiguousarray(arena, np.float64).copy()
    tmp = np.zeros(max(info["temp_elems"], 1))
    orc.run_streams(prog.streams(), [fac, tmp, None], info["n_insts"], info["scratch_elems"],
                    threads=threads)
    return fac


def residual(orc, hm, arena, x, b):
    """max over columns of ||b - A_H x|| / ||b|| with the oracle's numpy
    H-matrix-vector product."""
    r = b - orc.hmatvec(hm.leaves(), arena, x)
    return float(np.max(np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)))


def oracle_solution(pp, orc, hm, arena, b, tol, fac=None, threads=None, max_iter=60):
    """x = A_H^{-1} b by the CPU oracle in FP64: the oracle's factorization,
    substitution solve streams and iterative refinement with the oracle
    H-matvec until the residual is <= tol.  Returns (x, residual, sweeps)."""
    threads = threads or host_threads()
    if fac is None:
        fac = oracle_factor(pp, orc, hm, arena, threads)
    sprog = pp.Program.solve(hm, b.shape[1], threads)
    sinfo = sprog.info()

    def apply_inv(r):
        x = np.ascontiguousarray(r, np.float64).copy()
        orc.run_streams(sprog.streams(), [fac, None, x], sinfo["n_insts"], sinfo["scratch_elems"], threads=threads)
        return x

    x = apply_inv(b)
    it = 0
    while True:
        res = residual(orc, hm, arena, x, b)
        if res <= tol or it >= max_iter:
            return x, res, it
        x = x + apply_inv(b - orc.hmatvec(hm.leaves(), arena, x))
        it += 1
