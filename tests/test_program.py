"""Host pipeline: h-op generation, DAG, schedule, instruction encoding, and
the CPU oracle executing the encoded streams.

Pins:
  * FLOP model: the all-dense structure's instruction FLOPs sum to the
    reference's dense-LU closed form, whose value at N = 1024 is the
    spreadsheet's 715,303,424 (docs/SC/pspl results.xlsx!G2);
  * DAG: edges re-derived from the encoded words with an all-pairs
    checkOverlap (the reference's own build where available);
  * schedule: the program's per-worker queues equal the Python restatement
    of rts/scheduler.cpp run on the program's DAG;
  * numerics: the oracle executing the streams reproduces a dense no-pivot LU
    (all-dense case) and solves the H-matrix system (HODLR case).
"""
import numpy as np
import pytest

from oracle import scheduler as sref
from tests import wordmaps


def test_dense_flop_closed_form(pp):
    assert pp.dense_lu_flops(1024) == 715303424
    # N = 32768: 2.346e13 (docs/SC/sc-extended-abstract.tex:116)
    assert round(pp.dense_lu_flops(32768) / 1e13, 3) == 2.346
    for n, leaf in [(1024, 128), (512, 64), (768, 128)]:
        hm = pp.HMatrix("dense", n, leaf=leaf)
        assert pp.Program.factor(hm, 4).info()["flops"] == pp.dense_lu_flops(n)


def test_hodlr_structure(pp):
    hm = pp.HMatrix("hodlr", 16384, levels=6, rank=32)
    lv = hm.leaves()
    dense = [l for l in lv if l["type"] == "dense"]
    lowr = [l for l in lv if l["type"] == "lowrank"]
    assert len(dense) == 64 and all(l["ny"] == 256 and l["row0"] == l["col0"] for l in dense)
    assert len(lowr) == 2 * (2 ** 6 - 1) and all(l["rank"] == 32 for l in lowr)
    assert hm.storage >= 64 * 256 * 256 + sum((l["nx"] + l["ny"]) * 32 for l in lowr)
    info = pp.Program.factor(hm, 256).info()
    assert 1100 < info["n_insts"] < 1400
    assert 5.0e9 < info["flops"] < 6.5e9


def _decode_program(prog):
    n = prog.info()["n_insts"]
    words = [prog.inst_words(i) for i in range(n)]
    return n, words


@pytest.mark.parametrize("kind", ["dense", "hodlr"])
def test_dag_edges_all_pairs(pp, orc, kind):
    if kind == "dense":
        hm = pp.HMatrix("dense", 512, leaf=128)
    else:
        hm = pp.HMatrix("hodlr", 2048, levels=3, rank=8)
    prog = pp.Program.factor(hm, 4)
    n, words = _decode_program(prog)
    maps = [wordmaps.operand_maps(w) for w in words]
    if orc.ref_available():
        R = orc.ref()
        overlap = lambda a, b: bool(R.ref_check_overlap(*a, *b))
    else:
        overlap = pp.check_overlap
    for j in range(n):
        want = []
        for i in range(j):
            mask = 0
            for ro_f, mf in maps[i]:
                for ro_t, mt in maps[j]:
                    lo_f, hi_f = mf[0], mf[0] + mf[1] * mf[3]
                    lo_t, hi_t = mt[0], mt[0] + mt[1] * mt[3]
                    if hi_f <= lo_t or hi_t <= lo_f or not overlap(mf, mt):
                        continue
                    if not ro_f and ro_t:
                        mask |= 1
                    if ro_f and not ro_t:
                        mask |= 2
                    if not ro_f and not ro_t:
                        mask |= 4
            if mask:
                want.append((i, mask))
        assert prog.preds(j) == want, f"instruction {j}"


@pytest.mark.parametrize("workers", [1, 3, 16, 256])
def test_program_schedule_bit_exact(pp, workers):
    hm = pp.HMatrix("hodlr", 8192, levels=5, rank=16)
    prog = pp.Program.factor(hm, workers)
    n = prog.info()["n_insts"]
    pred = [prog.preds(j) for j in range(n)]
    succ = [[] for _ in range(n)]
    for j in range(n):
        for i, m in pred[j]:
            succ[i].append((j, m))
    flops = [prog.inst_flops(i) for i in range(n)]
    want = sref.schedule(n, flops, succ, pred, workers, pp.min_inst_flops())
    got = [prog.schedule(w) for w in range(workers)]
    assert got == want


def test_streams_encode_schedule(pp):
    hm = pp.HMatrix("hodlr", 4096, levels=4, rank=16)
    W = 8
    prog = pp.Program.factor(hm, W)
    n = prog.info()["n_insts"]
    seen = []
    for w, stream in enumerate(prog.streams()):
        ents = wordmaps.split_stream(stream)
        assert ents[-1][0] == "finish"
        sched = prog.schedule(w)
        body = ents[:-1]
        assert len(body) == len(sched)
        for (kind, words), (inst, ex, _) in zip(body, sched):
            assert (kind == "exec") == ex and int(words[1]) == inst
            if ex:
                assert list(words) == list(prog.inst_words(inst))
                seen.append(inst)
            else:
                assert 0 <= int(words[2]) <= 63
    assert sorted(seen) == list(range(n))


def test_oracle_executes_dense_lu(pp, orc):
    """All-dense H-LU through encoded streams == dense no-pivot LU."""
    n = 512
    hm, kern, arena = pp.build_config(0, n, leaf=128)
    a = pp.densify(hm, arena)
    prog = pp.Program.factor(hm, 3)
    fac = arena.copy()
    orc.run_streams(prog.streams(), [fac, None, None], prog.info()["n_insts"], 1 << 16, threads=1)
    lu_h = pp.densify(hm, fac)
    lu_d = orc.dense_lu(a)
    assert np.max(np.abs(lu_h - lu_d)) / np.max(np.abs(lu_d)) < 1e-12
    l, u = pp.lu_factors(lu_h)
    assert np.linalg.norm(a - l @ u) / np.linalg.norm(a) < 1e-14


def test_oracle_threaded_matches_sequential(pp, orc):
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    prog = pp.Program.factor(hm, 4)
    info = prog.info()
    f1, f2 = arena.copy(), arena.copy()
    t1, t2 = np.zeros(max(info["temp_elems"], 1)), np.zeros(max(info["temp_elems"], 1))
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)
    orc.run_streams(prog.streams(), [f1, t1, None], info["n_insts"], info["scratch_elems"], rnd=rnd, threads=1)
    orc.run_streams(prog.streams(), [f2, t2, None], info["n_insts"], info["scratch_elems"], rnd=rnd, threads=4)
    assert np.array_equal(f1, f2)


def test_oracle_hodlr_factor_and_solve(pp, orc):
    """HODLR H-LU + H-solve through the encoded streams: the solve residual
    against the densified A_H is at the level of the recompression error, and
    FP64 refinement with the H-LU factors drives it to 1e-12."""
    n = 2048
    hm, kern, arena = pp.build_config(1, n, levels=3, rank=32)
    a = pp.densify(hm, arena)
    prog = pp.Program.factor(hm, 4)
    info = prog.info()
    fac = arena.copy()
    tmp = np.zeros(max(info["temp_elems"], 1))
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)
    orc.run_streams(prog.streams(), [fac, tmp, None], info["n_insts"], info["scratch_elems"], rnd=rnd)
    nrhs = 4
    b = pp.rhs(1, n, nrhs)
    sprog = pp.Program.solve(hm, nrhs, 4)

    def apply_inv(r):
        x = np.ascontiguousarray(r.copy())
        orc.run_streams(sprog.streams(), [fac, None, x], sprog.info()["n_insts"], sprog.info()["scratch_elems"])
        return x

    x = apply_inv(b)
    res0 = orc.dense_residual(a, x, b)
    assert res0 < 1e-3
    for _ in range(12):
        r = b - a @ x
        if np.max(np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)) < 1e-12:
            break
        x = x + apply_inv(r)
    assert orc.dense_residual(a, x, b) < 1e-12


def test_oracle_trmm_matches_numpy(orc):
    """B <- T B of the oracle against numpy's product."""
    rng = np.random.default_rng(11)
    for ny, nx in [(1, 1), (7, 3), (64, 16), (33, 1)]:
        t = rng.standard_normal((ny, ny))
        b = rng.standard_normal((ny, nx))
        assert np.allclose(orc.trmm(b, t), t @ b, rtol=1e-13, atol=1e-13)


def test_oracle_inverse_factor_and_solve(pp, orc):
    """The factor program with leaf inversions leaves the same factors as the
    plain one (bit for bit), the inverse area holds L^{-1} and U^{-1} of every
    diagonal leaf, and the solve through the inverses agrees with the
    substitution solve and refines to 1e-12."""
    n = 2048
    hm, kern, arena = pp.build_config(1, n, levels=3, rank=32)
    a = pp.densify(hm, arena)
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)

    prog = pp.Program.factor(hm, 4)
    info = prog.info()
    fac = arena.copy()
    orc.run_streams(prog.streams(), [fac, np.zeros(max(info["temp_elems"], 1)), None], info["n_insts"],
                    info["scratch_elems"], rnd=rnd)

    iprog = pp.Program.factor_inv_batch([hm], 4)
    iinfo = iprog.info()
    ifac = hm.with_inverse_area(arena)
    orc.run_streams(iprog.streams(), [ifac, np.zeros(max(iinfo["temp_elems"], 1)), None], iinfo["n_insts"],
                    iinfo["scratch_elems"], rnd=rnd)
    assert np.array_equal(ifac[:hm.storage], fac[:hm.storage])

    base = hm.inverse_base
    cursor = 0
    for lf in hm.leaves():
        if lf["type"] != "dense" or lf["row0"] != lf["col0"]:
            continue
        nb = lf["ny"]
        lu = fac[lf["off"]:lf["off"] + nb * nb].reshape(nb, nb)
        l, u = np.tril(lu, -1) + np.eye(nb), np.triu(lu)
        second = (nb * nb + 31) // 32 * 32
        linv = ifac[base + cursor:base + cursor + nb * nb].reshape(nb, nb)
        uinv = ifac[base + cursor + second:base + cursor + second + nb * nb].reshape(nb, nb)
        assert np.max(np.abs(linv @ l - np.eye(nb))) < 1e-12
        assert np.max(np.abs(uinv @ u - np.eye(nb))) < 1e-10
        cursor += 2 * second
    leaves, nodes = hm.inverse_layout()
    assert cursor == (nodes[0]["l_u"] if nodes else ifac.size - base)

    nrhs = 4
    b = pp.rhs(1, n, nrhs)
    sprog = pp.Program.solve(hm, nrhs, 4)
    isprog = pp.Program.solve_inv_batch([hm], [nrhs], 4)
    kinds = {int(isprog.inst_words(i)[2]) for i in range(isprog.info()["n_insts"])}
    assert 10 in kinds and 2 not in kinds and 9 not in kinds  # trmm replaces trsml / trsmu

    def run(p, f, r):
        x = np.ascontiguousarray(r.copy())
        orc.run_streams(p.streams(), [f, None, x], p.info()["n_insts"], p.info()["scratch_elems"])
        return x

    x_sub, x_inv = run(sprog, fac, b), run(isprog, ifac, b)
    assert np.linalg.norm(x_inv - x_sub) / np.linalg.norm(x_sub) < 1e-10
    x = x_inv
    for _ in range(12):
        r = b - a @ x
        if np.max(np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)) < 1e-12:
            break
        x = x + run(isprog, ifac, r)
    assert orc.dense_residual(a, x, b) < 1e-12


def test_oracle_hmatvec_matches_dense(pp, orc):
    """The numpy H-matvec the CPU baseline refines with equals the densified
    matrix times x."""
    hm, kern, arena = pp.build_config(1, 1024, levels=2, rank=16)
    x = pp.rhs(1, 1024, 3)
    want = pp.densify(hm, arena) @ x
    got = orc.hmatvec(hm.leaves(), arena, x)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def dense_inverses(hm, area):
    """Dense L^{-1}, U^{-1} assembled from a filled inverse area."""
    leaves, nodes = hm.inverse_layout()
    n = hm.n
    li, ui = np.zeros((n, n)), np.zeros((n, n))
    for e in leaves:
        r, nb, o = e["row0"], e["nb"], e["inv_off"]
        second = (nb * nb + 31) // 32 * 32
        li[r:r + nb, r:r + nb] = area[o:o + nb * nb].reshape(nb, nb)
        ui[r:r + nb, r:r + nb] = area[o + second:o + second + nb * nb].reshape(nb, nb)
    for d in nodes:
        a0, na, nb, rl, ru = d["row0"], d["na"], d["nb"], d["rl"], d["ru"]
        b0 = a0 + na
        lu = area[d["l_u"]:d["l_u"] + nb * rl].reshape(nb, rl)
        lv = area[d["l_v"]:d["l_v"] + na * rl].reshape(na, rl)
        uu = area[d["u_u"]:d["u_u"] + na * ru].reshape(na, ru)
        uv = area[d["u_v"]:d["u_v"] + nb * ru].reshape(nb, ru)
        li[b0:b0 + nb, a0:a0 + na] = -lu @ lv.T
        ui[a0:a0 + na, b0:b0 + nb] = -uu @ uv.T
    return li, ui


@pytest.mark.parametrize("n,levels,rank", [(1024, 2, 16), (2048, 3, 32)])
def test_oracle_hierarchical_inverse(pp, orc, n, levels, rank):
    """The factor program with the hierarchical inverse leaves factors equal
    (bit for bit) to the plain program's, and the inverse area assembles to
    numpy's inverses of the dense unit-lower and upper factors."""
    hm, kern, arena = pp.build_config(1, n, levels=levels, rank=rank)
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)
    prog = pp.Program.factor(hm, 4)
    info = prog.info()
    fac = arena.copy()
    orc.run_streams(prog.streams(), [fac, np.zeros(max(info["temp_elems"], 1)), None], info["n_insts"],
                    info["scratch_elems"], rnd=rnd)
    iprog = pp.Program.factor_inv_batch([hm], 4)
    iinfo = iprog.info()
    assert iinfo["n_insts"] > info["n_insts"] + 2 * len(hm.inverse_layout()[0])
    ifac = hm.with_inverse_area(arena)
    orc.run_streams(iprog.streams(), [ifac, np.zeros(max(iinfo["temp_elems"], 1)), None], iinfo["n_insts"],
                    iinfo["scratch_elems"], rnd=rnd)
    assert np.array_equal(ifac[:hm.storage], fac[:hm.storage])
    l, u = pp.lu_factors(pp.densify(hm, fac))
    li, ui = dense_inverses(hm, ifac[hm.inverse_base:])
    assert np.max(np.abs(li - np.linalg.inv(l))) < 1e-11 * np.max(np.abs(np.linalg.inv(l)))
    assert np.max(np.abs(ui - np.linalg.inv(u))) < 1e-11 * np.max(np.abs(np.linalg.inv(u)))


def test_instruction_bytes(pp):
    """Algorithmic bytes per instruction: operands read once, written twice."""
    hm = pp.HMatrix("dense", 512, leaf=128)
    prog = pp.Program.factor(hm, 2)
    seen = set()
    for i in range(prog.info()["n_insts"]):
        w = prog.inst_words(i)
        op = int(w[2])
        b = prog.inst_bytes(i)
        if op == 1:  # getrf of a 128 x 128 leaf: read and written
            assert b == 2 * 128 * 128 * 8
        elif op in (2, 3):  # trsm: 128 x 128 B written, the triangular factor read
            assert b == 2 * 128 * 128 * 8 + 128 * 128 * 8
        elif op == 4:  # gemm M -= A B on 128 x 128 blocks
            assert b == 2 * 128 * 128 * 8 + 2 * 128 * 128 * 8
        seen.add(op)
    assert {1, 2, 3, 4} <= seen
