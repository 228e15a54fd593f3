"""Byte-string format on the GPU.

  * factors saved from one solver and loaded into a fresh solver over the
    same matrix solve to bit-identical results without factoring again (every
    solve mode, FP64 and mixed precision);
  * a string from another precision, solve mode or matrix entries is
    rejected;
  * a decoded factor program runs on the VM and factors bit-identically to
    the program it was encoded from.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solve(solver, db, nrhs):
    dx = torch.zeros_like(db)
    rep = solver.solve(db.data_ptr(), dx.data_ptr(), nrhs, tol=1e-10, max_iter=40)
    torch.cuda.synchronize()
    return dx.cpu().numpy(), rep


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_factor_strings_round_trip(pp, ctx, precision, mode):
    n, nrhs = 2048, 4
    print(f"[case] factors n={n} precision={precision} mode={mode}", flush=True)
    hm, kern, arena = pp.build_config(1, n, levels=3, rank=16)
    b = pp.rhs(1, n, nrhs)
    db = torch.from_numpy(b).cuda()
    s1 = pp.Solver.batch(ctx, [hm], [arena], precision=precision, workers=32, nrhs_max=nrhs, solve_mode=mode)
    s1.factor()
    x1, rep1 = _solve(s1, db, nrhs)
    data = s1.factors_to_bytes()
    s2 = pp.Solver.batch(ctx, [hm], [arena], precision=precision, workers=32, nrhs_max=nrhs, solve_mode=mode)
    s2.factors_from_bytes(data)
    x2, rep2 = _solve(s2, db, nrhs)
    assert rep1["converged"] and rep2["converged"]
    assert rep2["iterations"] == rep1["iterations"]
    assert np.array_equal(x1, x2)
    assert s2.factors_to_bytes() == data
    s1.close()
    s2.close()


def test_factor_strings_rejected_elsewhere(pp, ctx):
    n, nrhs = 1024, 2
    print(f"[case] factor string checks n={n}", flush=True)
    hm, kern, arena = pp.build_config(1, n, levels=2, rank=16)
    s1 = pp.Solver.batch(ctx, [hm], [arena], precision=0, workers=16, nrhs_max=nrhs, solve_mode=0)
    s1.factor()
    data = s1.factors_to_bytes()
    other_prec = pp.Solver.batch(ctx, [hm], [arena], precision=2, workers=16, nrhs_max=nrhs, solve_mode=0)
    other_mode = pp.Solver.batch(ctx, [hm], [arena], precision=0, workers=16, nrhs_max=nrhs, solve_mode=2)
    a2 = arena.copy()
    a2[7] += 1.0
    other_entries = pp.Solver.batch(ctx, [hm], [a2], precision=0, workers=16, nrhs_max=nrhs, solve_mode=0)
    for s, what in ((other_prec, "layout"), (other_mode, "layout"), (other_entries, "entries")):
        with pytest.raises(pp.PastelError) as e:
            s.factors_from_bytes(data)
        assert what in str(e.value), str(e.value)
        s.close()
    bad = bytearray(data)
    bad[len(bad) // 2] ^= 1
    with pytest.raises(pp.PastelError):
        s1.factors_from_bytes(bytes(bad))
    s1.close()


def test_decoded_program_runs(pp, orc, ctx):
    print("[case] decoded factor program n=2048", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    prog = pp.Program.factor(hm, 16)
    dec = pp.Program.from_bytes(prog.to_bytes())
    out = []
    for p in (prog, dec):
        d = torch.from_numpy(arena).cuda()
        s = pp.Stream(ctx, p, [d.data_ptr(), None, None], 0)
        s.set_watchdog(10000)
        s.launch(4)
        s.wait()
        torch.cuda.synchronize()
        out.append(d.cpu().numpy())
        s.close()
    assert np.array_equal(out[0], out[1])
    from tests import hlu_host as H
    worst = H.worst_leaf_diff(hm, out[1], H.oracle_factor(pp, orc, hm, arena))
    print(f"decoded program vs oracle: worst leaf difference {worst:.2e}")
    assert worst < 1e-9
