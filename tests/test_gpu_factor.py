"""GPU H-LU factor + solve through the persistent VM, against the CPU oracle.

  * the GPU-factored arena equals the oracle executing the same encoded
    streams (FP64: 1e-12 relative on every dense leaf and low-rank product
    for the all-dense case, 1e-9 for HODLR where accum's projection adds
    conditioning);
  * all-dense factors reproduce the dense no-pivot LU oracle;
  * solve + FP64 refinement reach a relative residual <= 1e-10 against the
    densified A_H (oracle dense residual), FP64 and mixed precision;
  * the VM watchdog turns a lost signal into an error, not a hang.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def leaf_products(hm, arena):
    out = []
    for lf in hm.leaves():
        ny, nx = lf["ny"], lf["nx"]
        if lf["type"] == "dense":
            out.append(arena[lf["off"]:lf["off"] + ny * nx].reshape(ny, nx))
        else:
            r = lf["rank"]
            u = arena[lf["off"]:lf["off"] + ny * r].reshape(ny, r)
            v = arena[lf["off_v"]:lf["off_v"] + nx * r].reshape(nx, r)
            out.append(u @ v.T)
    return out


def gpu_factor(pp, ctx, hm, arena, workers, precision=0, wavefronts=4):
    prog = pp.Program.factor(hm, workers)
    dt = torch.float64 if precision == 0 else torch.float32
    d = torch.from_numpy(arena).to(dt).cuda()
    s = pp.Stream(ctx, prog, [d.data_ptr(), None, None], precision)
    s.set_watchdog(10000)
    s.launch(wavefronts)
    s.wait()
    torch.cuda.synchronize()
    return prog, d.double().cpu().numpy(), s


def oracle_factor(pp, orc, prog, arena):
    info = prog.info()
    fac = arena.copy()
    tmp = np.zeros(max(info["temp_elems"], 1))
    orc.run_streams(prog.streams(), [fac, tmp, None], info["n_insts"], info["scratch_elems"])
    return fac


@pytest.mark.parametrize("workers", [1, 8, 64])
def test_dense_factor_parity(pp, orc, ctx, workers):
    print(f"[case] dense n=1024 leaf=128 workers={workers}", flush=True)
    hm, kern, arena = pp.build_config(0, 1024)
    prog, gfac, _ = gpu_factor(pp, ctx, hm, arena, workers)
    ofac = oracle_factor(pp, orc, prog, arena)
    for g, o in zip(leaf_products(hm, gfac), leaf_products(hm, ofac)):
        assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o))
    a = pp.densify(hm, arena)
    l, u = pp.lu_factors(pp.densify(hm, gfac))
    assert np.linalg.norm(a - l @ u) / np.linalg.norm(a) < 1e-14


@pytest.mark.parametrize("n,levels,rank,workers", [(2048, 3, 16, 8), (4096, 4, 32, 32)])
def test_hodlr_factor_parity(pp, orc, ctx, n, levels, rank, workers):
    print(f"[case] hodlr n={n} levels={levels} rank={rank} workers={workers}", flush=True)
    hm, kern, arena = pp.build_config(1, n, levels=levels, rank=rank)
    prog, gfac, _ = gpu_factor(pp, ctx, hm, arena, workers)
    ofac = oracle_factor(pp, orc, prog, arena)
    worst = 0.0
    for g, o in zip(leaf_products(hm, gfac), leaf_products(hm, ofac)):
        worst = max(worst, np.max(np.abs(g - o)) / max(np.max(np.abs(o)), 1e-300))
    print("worst leaf relative difference", worst)
    assert worst < 1e-9


def test_inverse_factor_parity(pp, orc, ctx):
    """Factor program with leaf inversions: the factors and the inverse area
    on the GPU equal the oracle running the same streams."""
    print("[case] inverse factor hodlr n=2048 levels=3 rank=16 workers=16", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    prog = pp.Program.factor_inv_batch([hm], 16)
    full = hm.with_inverse_area(arena)
    d = torch.from_numpy(full).cuda()
    s = pp.Stream(ctx, prog, [d.data_ptr(), None, None], 0)
    s.set_watchdog(10000)
    s.launch(4)
    s.wait()
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    want = oracle_factor(pp, orc, prog, full)
    for g, o in zip(leaf_products(hm, got), leaf_products(hm, want)):
        assert np.max(np.abs(g - o)) <= 1e-9 * max(np.max(np.abs(o)), 1e-300)
    gi, oi = got[hm.inverse_base:], want[hm.inverse_base:]
    assert np.max(np.abs(gi - oi)) <= 1e-10 * np.max(np.abs(oi))
    s.close()


def test_factor_repeatable(pp, ctx):
    """Two launches of the same stream (epoch-tagged flags, no reset) give
    bit-identical factors."""
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    prog = pp.Program.factor(hm, 16)
    d = torch.from_numpy(arena).cuda()
    s = pp.Stream(ctx, prog, [d.data_ptr(), None, None], 0)
    outs = []
    for _ in range(3):
        d.copy_(torch.from_numpy(arena))
        s.launch(4)
        s.wait()
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy().copy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("cfg,n", [(0, 1024), (1, 4096)])
def test_solve_refined(pp, orc, ctx, precision, cfg, n):
    print(f"[case] solve cfg={cfg} n={n} precision={precision}", flush=True)
    if cfg == 0:
        hm, kern, arena = pp.build_config(0, n)
        nrhs = 1
    else:
        hm, kern, arena = pp.build_config(1, n, levels=4, rank=32)
        nrhs = 16
    solver = pp.Solver(ctx, hm, arena, precision=precision, workers=32, nrhs_max=nrhs)
    solver.factor()
    b = pp.rhs(cfg, n, nrhs)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-10, max_iter=40)
    x = dx.cpu().numpy()
    a = pp.densify(hm, arena)
    res = orc.dense_residual(a, x, b)
    print("report", rep, "oracle residual", res)
    assert rep["converged"] and rep["residual"] <= 1e-10
    assert res <= 2e-10
    # error against the FP64 dense-LU oracle solution
    xo = orc.dense_lu_solve(orc.dense_lu(a), b)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("error vs FP64 oracle", err)
    assert err < 1e-8


def test_padded_solve_widths(pp, ctx):
    """Requests with 3 and 11 columns run the 4- and 16-wide solve programs
    over zero-padded right-hand sides; each column equals the same column
    solved in a full 16-column request."""
    print("[case] padded solve widths n=2048", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    solver = pp.Solver(ctx, hm, arena, precision=2, workers=16, nrhs_max=16)
    solver.factor()
    b = pp.rhs(1, 2048, 16)
    full = torch.zeros(2048, 16, dtype=torch.float64, device="cuda")
    rep = solver.solve(torch.from_numpy(b).cuda().data_ptr(), full.data_ptr(), 16, tol=1e-12, max_iter=40)
    assert rep["converged"]
    want = full.cpu().numpy()
    for k in (3, 11):
        db = torch.from_numpy(np.ascontiguousarray(b[:, :k])).cuda()
        dx = torch.zeros_like(db)
        r = solver.solve(db.data_ptr(), dx.data_ptr(), k, tol=1e-12, max_iter=40)
        assert r["converged"]
        got = dx.cpu().numpy()
        assert np.linalg.norm(got - want[:, :k]) <= 1e-10 * np.linalg.norm(want[:, :k])


@pytest.mark.parametrize("nrhs", [4, 24, 80])
def test_matvec_matches_dense(pp, ctx, nrhs):
    """The FP64 residual product against numpy: one, two and (past 64
    columns) two groups of four 16-wide column tiles."""
    print(f"[case] matvec n=2048 nrhs={nrhs}", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    solver = pp.Solver(ctx, hm, arena, precision=0, workers=8, nrhs_max=nrhs)
    rng = np.random.default_rng(3 + nrhs)
    x = rng.standard_normal((2048, nrhs))
    b = rng.standard_normal((2048, nrhs))
    dx, db = torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda()
    dr = torch.zeros_like(db)
    solver.residual(db.data_ptr(), dx.data_ptr(), dr.data_ptr(), nrhs)
    want = b - pp.densify(hm, arena) @ x
    assert np.max(np.abs(dr.cpu().numpy() - want)) < 1e-10 * np.max(np.abs(want))


@pytest.mark.parametrize("precision", [0, 2])
def test_solve_wide(pp, orc, ctx, precision):
    """80 right-hand sides (two column groups in every H-matrix product and in
    the residual norms): refined to 1e-10, checked against the dense FP64
    oracle residual and solution."""
    print(f"[case] wide solve n=2048 nrhs=80 precision={precision}", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    nrhs = 80
    solver = pp.Solver(ctx, hm, arena, precision=precision, workers=16, nrhs_max=nrhs)
    solver.factor()
    b = np.random.default_rng(11).standard_normal((2048, nrhs))
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-10, max_iter=40)
    x = dx.cpu().numpy()
    a = pp.densify(hm, arena)
    res = orc.dense_residual(a, x, b)
    assert rep["converged"] and rep["residual"] <= 1e-10
    assert res <= 2e-10
    xo = orc.dense_lu_solve(orc.dense_lu(a), b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_watchdog_reports_lost_signal(pp, ctx):
    """A wait on an instruction no worker executes ends with a timeout error
    (the wait is bounded by a 200 ms watchdog set from the host)."""
    import ctypes as C
    L = pp.lib()
    words = [np.array([2, 0, 0, 0], np.int32), np.array([0], np.int32)]
    ptrs = (C.c_void_p * 1)(None)
    arr = (C.POINTER(C.c_int32) * 2)(*[w.ctypes.data_as(C.POINTER(C.c_int32)) for w in words])
    lens = np.array([4, 1], np.int32)
    h = C.c_void_p()
    pp.check(L.pp_stream_upload(ctx.h, arr, lens.ctypes.data_as(C.POINTER(C.c_int32)), 2, ptrs, 1, 1024, 1, 0,
                                C.byref(h)))
    pp.check(L.pp_stream_set_watchdog(ctx.h, h, 200))
    pp.check(L.pp_launch(ctx.h, h, 4, None))
    first = C.c_int32(-9)
    rc = L.pp_wait(ctx.h, h, 30000, C.byref(first))
    assert rc == -3 and first.value == 0
    L.pp_stream_destroy(ctx.h, h)


def test_upload_rejects_unknown_opcode(pp, ctx):
    import ctypes as C
    L = pp.lib()
    w = np.array([7, 0, 0], np.int32)
    arr = (C.POINTER(C.c_int32) * 1)(w.ctypes.data_as(C.POINTER(C.c_int32)))
    lens = np.array([3], np.int32)
    h = C.c_void_p()
    rc = L.pp_stream_upload(ctx.h, arr, lens.ctypes.data_as(C.POINTER(C.c_int32)), 1, (C.c_void_p * 1)(None), 1, 64,
                            1, 0, C.byref(h))
    assert rc == -1


def test_mixed_rank48_accum_without_fallback(pp, ctx, orc):
    """Mixed precision at rank 48: every accum orthonormalises by CholeskyQR
    (no Householder fallback), and the refined solve still reaches 1e-10."""
    n, nrhs = 4096, 4
    print(f"[case] mixed rank 48 n={n}", flush=True)
    hm, kern, arena = pp.build_config(1, n, levels=4, rank=48)
    prog = pp.Program.factor_inv_batch([hm], 64)
    full = hm.with_inverse_area(arena)
    d = torch.from_numpy(full).float().cuda()
    s = pp.Stream(ctx, prog, [d.data_ptr(), None, None], pp.MIXED)
    s.set_watchdog(20000)
    s.launch(4)
    s.wait()
    assert s.qr_fallbacks() == 0
    s.close()
    solver = pp.Solver(ctx, hm, arena, precision=pp.MIXED, workers=64, nrhs_max=nrhs)
    solver.factor()
    b = pp.rhs(1, n, nrhs)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-10, max_iter=40)
    solver.close()
    assert rep["converged"] and rep["residual"] <= 1e-10


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("n,levels,rank,nrhs", [(1000, 2, 20, 1), (3000, 3, 24, 3)])
def test_ragged_sizes(pp, orc, ctx, precision, n, levels, rank, nrhs):
    """Sizes that are no multiple of 32 (leaves of 250 and 375 rows, ranks 20
    and 24, one right-hand side): the refined solve reaches 1e-10 against the
    dense FP64 oracle of A_H and agrees with its dense LU solution."""
    print(f"[case] ragged n={n} levels={levels} rank={rank} nrhs={nrhs} precision={precision}", flush=True)
    hm, kern, arena = pp.build_config(1, n, levels=levels, rank=rank)
    solver = pp.Solver(ctx, hm, arena, precision=precision, workers=32, nrhs_max=nrhs)
    solver.factor()
    b = pp.rhs(2, n, nrhs)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-10, max_iter=40)
    solver.close()
    x = dx.cpu().numpy()
    a = pp.densify(hm, arena)
    assert rep["converged"]
    assert orc.dense_residual(a, x, b) < 1e-9
    xo = orc.dense_lu_solve(orc.dense_lu(a), b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7


def test_ragged_factor_parity(pp, orc, ctx):
    """The GPU factors of a ragged HODLR equal the oracle's on the same streams."""
    print("[case] ragged factor parity n=1000", flush=True)
    hm, kern, arena = pp.build_config(1, 1000, levels=2, rank=20)
    prog, gfac, _ = gpu_factor(pp, ctx, hm, arena, 8)
    ofac = oracle_factor(pp, orc, prog, arena)
    for g, o in zip(leaf_products(hm, gfac), leaf_products(hm, ofac)):
        assert np.linalg.norm(g - o) / max(np.linalg.norm(o), 1e-300) < 1e-9
