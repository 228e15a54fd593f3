"""Windowed factorization (src/launcher.cu:146-167): the h-op list cut into
consecutive windows, each with its own DAG and schedule (CPU).

  * the windows hold the factorization's instructions in order, and each
    window's dependency edges are the whole DAG's edges inside it;
  * the oracle running the windows one after the other, with one temp area
    cleared only before the first, factors bit-identically to the whole
    program;
  * the temp carry-over flag survives the byte-string format.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("window", [7, 50, 0])
def test_windows_partition_the_program(pp, window):
    hm = pp.HMatrix("hodlr", 2048, levels=3, rank=16)
    full = pp.Program.factor(hm, 4)
    n = full.info()["n_insts"]
    wins = pp.Program.factor_windows(hm, 4, window)
    assert len(wins) == (1 if window <= 0 else -(-n // window))
    at = 0
    for w in wins:
        m = w.info()["n_insts"]
        for i in range(m):
            a, b = w.inst_words(i), full.inst_words(at + i)
            assert a[1] == i and b[1] == at + i  # the id slot holds the window-local id
            assert np.array_equal(np.delete(a, 1), np.delete(b, 1))
            inside = sorted((p - at, k) for p, k in full.preds(at + i) if p >= at)
            assert sorted(w.preds(i)) == inside
        assert w.temp_elems(0) == full.temp_elems(0)
        at += m
    assert at == n


@pytest.mark.parametrize("window", [5, 40])
def test_windows_factor_like_the_whole_program(pp, orc, window):
    hm, kern, arena = pp.build_config(1, 2048, levels=3, rank=16)
    full = pp.Program.factor(hm, 4)
    info = full.info()
    rnd = pp.sketch(pp.SKETCH_SEED, 1 << 16)
    ref = arena.copy()
    tmp = np.zeros(max(info["temp_elems"], 1))
    orc.run_streams(full.streams(), [ref, tmp, None], info["n_insts"], info["scratch_elems"], rnd=rnd, threads=1)
    fac = arena.copy()
    tmp = np.zeros(max(info["temp_elems"], 1))  # cleared once, before the first window
    for w in pp.Program.factor_windows(hm, 4, window):
        wi = w.info()
        orc.run_streams(w.streams(), [fac, tmp, None], wi["n_insts"], max(wi["scratch_elems"], 1), rnd=rnd,
                        threads=1)
    assert np.array_equal(fac, ref)


def test_window_flag_survives_the_format(pp):
    hm = pp.HMatrix("hodlr", 1024, levels=2, rank=8)
    wins = pp.Program.factor_windows(hm, 2, 6)
    assert len(wins) > 1
    for i, w in enumerate(wins):
        data = w.to_bytes()
        assert pp.Program.from_bytes(data).to_bytes() == data
        flags = int.from_bytes(data[16 + 16 + 16:16 + 16 + 20], "little")
        assert flags == (0 if i == 0 else 1)
