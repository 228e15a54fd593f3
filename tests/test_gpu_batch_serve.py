"""Batched execution (many H-matrices in one persistent launch, replayed from a
captured hipGraph) and the serving loop, on the GPU.

  * a batch gives each member the solution of the dense FP64 oracle (LU of
    the densified member, 1e-10 relative) and the same solution as a solver
    over that member alone (tolerance 1e-10 relative after both are refined
    to 1e-12);
  * graph replays give bit-identical factors to eager launches, with eager
    work between replays, and every member's factors equal the oracle's
    execution of its own factorization program (leaf products, 1e-9);
  * serving: every request converges to the dense FP64 oracle's solution,
    and results are identical whether the request runs alone or
    concurrently with others.
"""
import numpy as np
import pytest
import torch

from tests import hlu_host as H

pytestmark = pytest.mark.gpu


def _members(pp, sizes, rank=16):
    out = []
    for i, n in enumerate(sizes):
        levels = max(1, int(np.log2(n // 256)))
        hm = pp.HMatrix("hodlr", n, levels=levels, rank=rank)
        k = pp.Kernel(pp.points_2d_morton(3 + i, n), "inv2d", n / 4.0)
        out.append((hm, pp.fill_arena(hm, k)))
    return out


@pytest.mark.parametrize("precision", [0, 2])
def test_batch_matches_single(pp, orc, ctx, precision):
    print(f"[case] batch precision={precision}", flush=True)
    mem = _members(pp, [512, 1024, 768, 2048])
    nrhs = 4
    bs = [pp.rhs(3, hm.n, nrhs) for hm, _ in mem]
    batch = pp.Solver.batch(ctx, [m[0] for m in mem], [m[1] for m in mem], precision=precision, workers=32,
                            nrhs_max=nrhs)
    batch.factor()
    B = np.concatenate(bs, axis=0)
    db = torch.from_numpy(B).cuda()
    dx = torch.zeros_like(db)
    rep = batch.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-12, max_iter=60)
    assert rep["converged"]
    X = dx.cpu().numpy()
    for (hm, arena), b, m in zip(mem, bs, batch.members()):
        s = pp.Solver(ctx, hm, arena, precision=precision, workers=32, nrhs_max=nrhs)
        s.factor()
        d1 = torch.from_numpy(b).cuda()
        x1 = torch.zeros_like(d1)
        r1 = s.solve(d1.data_ptr(), x1.data_ptr(), nrhs, tol=1e-12, max_iter=60)
        assert r1["converged"] and m["residual"] <= 1e-12
        xb = X[m["row_off"]:m["row_off"] + hm.n]
        xs = x1.cpu().numpy()
        assert np.linalg.norm(xb - xs) / np.linalg.norm(xs) < 1e-10
        a = pp.densify(hm, arena)
        xo = orc.dense_lu_solve(orc.dense_lu(a), b)
        assert np.linalg.norm(xb - xo) / np.linalg.norm(xo) < 1e-10
        assert orc.dense_residual(a, xb, b) < 2e-12
        s.close()
    batch.close()


def test_graph_replay_matches_eager(pp, orc, ctx):
    mem = _members(pp, [1024, 512, 1536])
    s = pp.Solver.batch(ctx, [m[0] for m in mem], [m[1] for m in mem], precision=0, workers=16, nrhs_max=2)
    s.factor()
    ptr, eb = s.factor_arena()
    n_elem = int(pp.lib().pp_solver_storage_elems(s.h))  # every member's arena and inverse area
    view = torch.empty(n_elem, dtype=torch.float64, device="cuda")

    def snapshot():
        torch.cuda.synchronize()
        import ctypes as C
        out = np.empty(n_elem)
        pp.check(pp.lib().pp_copy_d2h(ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), 8 * n_elem))
        return out

    eager = snapshot()
    s.use_graph(True)
    stream = torch.cuda.Stream()
    for rep in range(4):
        s.factor(stream.cuda_stream)
        got = snapshot()
        assert np.array_equal(got, eager), f"replay {rep}"
        view.normal_()  # eager work between replays
        torch.cuda.synchronize()
    s.use_graph(False)
    s.factor()
    assert np.array_equal(snapshot(), eager)
    for (hm, arena), m in zip(mem, s.members()):
        worst = H.worst_leaf_diff(hm, eager, H.oracle_factor(pp, orc, hm, arena), base_got=m["arena_off"])
        assert worst < 1e-9, (hm.n, worst)
    s.close()


def test_serving_alone_equals_mix(pp, orc, ctx):
    from importlib import import_module
    serve = import_module("pastel-palettes_amd.serve")
    sizes = {"a": 1024, "b": 2048, "c": 1536}
    mats = {}
    for i, (key, n) in enumerate(sizes.items()):
        levels = max(1, int(np.log2(n // 256)))
        hm = pp.HMatrix("hodlr", n, levels=levels, rank=16)
        arena = pp.fill_arena(hm, pp.Kernel(pp.points_2d_morton(10 + i, n), "inv2d", n / 4.0))
        mats[key] = (hm, torch.from_numpy(arena).cuda())

    def make_rhs(key, nrhs, k):
        return np.ascontiguousarray(pp.rhs(4 + k, sizes[key], nrhs))

    trace = [(0.0, "a", 3, False), (0.0, "b", 5, False), (0.0, "c", 1, False), (0.0, "a", 2, True),
             (0.0, "b", 7, True), (0.0, "c", 4, True)]
    # alone: one slot, requests one after another
    s1 = serve.Server(pp, torch, ctx, mats, pp.MIXED, budget_bytes=8 << 30, slots=1, workers_per_slot=32)
    alone, _ = serve.replay(s1, trace, make_rhs, realtime=False)
    s1.close()
    # mix: three slots, all at once
    s3 = serve.Server(pp, torch, ctx, mats, pp.MIXED, budget_bytes=8 << 30, slots=3, workers_per_slot=32)
    mix, _ = serve.replay(s3, trace, make_rhs, realtime=False)
    s3.close()
    for a, m in zip(alone, mix):
        assert a.error is None and m.error is None, (a.error, m.error)
        assert a.report["converged"] and m.report["converged"]
        assert a.report["residual"] <= 1e-10
        assert np.array_equal(a.x, m.x), f"request {a.rid} differs alone vs in the mix"
        hm, arena = mats[m.key]
        dense = pp.densify(hm, arena.cpu().numpy())
        xo = orc.dense_lu_solve(orc.dense_lu(dense), m.b)
        assert np.linalg.norm(m.x - xo) / np.linalg.norm(xo) < 1e-8
        assert orc.dense_residual(dense, m.x, m.b) <= 2e-10


@pytest.mark.parametrize("mode", [1, 2])
def test_solve_modes_agree(pp, ctx, mode):
    """Leaf-inverse and substitution solves converge to the solution of the
    hierarchical-inverse solve (tolerance 1e-10 relative after refinement to
    1e-12)."""
    print(f"[case] solve mode {mode}", flush=True)
    mem = _members(pp, [512, 1024, 768])
    nrhs = 3
    B = np.concatenate([pp.rhs(3, hm.n, nrhs) for hm, _ in mem], axis=0)
    out = {}
    for m in (0, mode):
        s = pp.Solver.batch(ctx, [x[0] for x in mem], [x[1] for x in mem], precision=2, workers=16, nrhs_max=4,
                            solve_mode=m)
        s.factor()
        db = torch.from_numpy(B).cuda()
        dx = torch.zeros_like(db)
        rep = s.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-12, max_iter=60)
        assert rep["converged"]
        out[m] = dx.cpu().numpy()
        s.close()
    assert np.linalg.norm(out[mode] - out[0]) <= 1e-10 * np.linalg.norm(out[0])
