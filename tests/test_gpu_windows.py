"""Windowed factorization on the GPU: the windows launched one after the
other on one shared temp buffer factor bit-identically to the single launch
(FP64 and mixed precision), and both equal the CPU oracle's execution of the
whole program (FP64: leaf products within 1e-9 relative; mixed: within the
BF16 factor tolerance of the configuration tests); a later window without
the shared buffer is refused at upload."""
import numpy as np
import pytest
import torch

from tests import hlu_host as H
from tests.test_gpu_configs import MIXED_FACTOR_TOL

pytestmark = pytest.mark.gpu


def run(pp, ctx, progs, arena, precision):
    dt = torch.float64 if precision == pp.FP64 else torch.float32
    d = torch.from_numpy(arena).to(dt).cuda()
    tmp = torch.zeros(max(progs[0].temp_elems(0), 1), dtype=dt, device="cuda")
    for p in progs:
        s = pp.Stream(ctx, p, [d.data_ptr(), tmp.data_ptr(), None], precision)
        s.set_watchdog(10000)
        s.launch(4)
        s.wait()
        s.close()
    torch.cuda.synchronize()
    return d.double().cpu().numpy()


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("window", [9, 64])
def test_windows_match_single_launch(pp, orc, ctx, precision, window):
    print(f"[case] windows n=2048 window={window} precision={precision}", flush=True)
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    whole = run(pp, ctx, [pp.Program.factor(hm, 16)], arena, precision)
    wins = pp.Program.factor_windows(hm, 16, window)
    assert len(wins) > 1
    got = run(pp, ctx, wins, arena, precision)
    assert np.array_equal(got, whole)
    worst = H.worst_leaf_diff(hm, got, H.oracle_factor(pp, orc, hm, arena))
    print(f"windows vs oracle: worst leaf difference {worst:.2e}")
    assert worst < (1e-9 if precision == pp.FP64 else MIXED_FACTOR_TOL)


def test_later_window_needs_the_shared_temps(pp, ctx):
    print("[case] window upload without temps", flush=True)
    hm, kern, arena = pp.build_config(1, 1024, levels=2, rank=8)
    wins = pp.Program.factor_windows(hm, 8, 10)
    d = torch.from_numpy(arena).cuda()
    with pytest.raises(pp.PastelError) as e:
        pp.Stream(ctx, wins[1], [d.data_ptr(), None, None], pp.FP64)
    assert "temp buffer" in str(e.value)
