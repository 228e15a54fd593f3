"""Drop-in rts layer: overlap test, dependency classification, scheduler.

Pins:
  * overlap: brute-force byte scan (the method of the reference's main.cpp)
    and, where /root/reference exists, the reference's own compiled
    checkOverlap (oracle/_ref);
  * dependency masks: the truth table of rts/ops.cpp:50-57;
  * scheduler: two schedules traced by hand from rts/scheduler.cpp pin the
    Python restatement (oracle/scheduler.py), which then checks the C++
    scheduler bit-exactly on random DAGs and on the real H-LU DAGs.
"""
import numpy as np
import pytest

from oracle import scheduler as sref


def brute_overlap(a, b):
    """Byte-level scan: is some byte in both strided regions?"""
    def bytes_of(m):
        p, pitch, w, h = m
        s = set()
        for r in range(h):
            s.update(range(p + r * pitch, p + r * pitch + w))
        return s
    return len(bytes_of(a) & bytes_of(b)) > 0


def random_maps(rng, count, allow_wide=False):
    out = []
    for _ in range(count):
        a, b = rng.integers(0, 200, 2)
        ma, na, oa = rng.integers(1, 17), rng.integers(1, 17), rng.integers(0, 16)
        mb, nb, ob = rng.integers(1, 17), rng.integers(1, 17), rng.integers(0, 16)
        out.append(((int(a), int(ma + oa), int(ma), int(na)), (int(b), int(mb + ob), int(mb), int(nb))))
    return out


def test_overlap_brute_force(pp):
    rng = np.random.default_rng(12345)
    for a, b in random_maps(rng, 2000):
        assert pp.check_overlap(a, b) == brute_overlap(a, b), (a, b)


def test_overlap_edge_cases(pp):
    # identical start, touching rows, adjacent columns, one row
    assert pp.check_overlap((100, 8, 8, 4), (100, 8, 8, 4))
    assert not pp.check_overlap((0, 16, 8, 4), (8, 16, 8, 4))      # interleaved columns
    assert pp.check_overlap((0, 16, 9, 4), (8, 16, 8, 4))          # one byte shared
    assert not pp.check_overlap((0, 16, 16, 4), (64, 16, 16, 4))   # back to back
    assert pp.check_overlap((0, 16, 16, 4), (63, 16, 1, 1))
    assert not pp.check_overlap((0, 1 << 20, 8, 1 << 10), (8, 1 << 20, 8, 1 << 10))


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="reference rts library not built here")
def test_overlap_matches_reference_build(pp, orc):
    R = orc.ref()
    rng = np.random.default_rng(777)
    for a, b in random_maps(rng, 5000):
        assert pp.check_overlap(a, b) == bool(R.ref_check_overlap(*a, *b)), (a, b)
    # degenerate regions (zero width / height) follow the reference too
    for _ in range(2000):
        a = (int(rng.integers(0, 64)), int(rng.integers(1, 9)), int(rng.integers(0, 9)), int(rng.integers(0, 4)))
        b = (int(rng.integers(0, 64)), int(rng.integers(1, 9)), int(rng.integers(0, 9)), int(rng.integers(0, 4)))
        assert pp.check_overlap(a, b) == bool(R.ref_check_overlap(*a, *b)), (a, b)


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="reference rts library not built here")
def test_reference_dependency_table_semantics(orc):
    """The reference container keeps the first insertion order per key and
    removes an edge on a `no` update (rts/dependency.cpp:11-34)."""
    R = orc.ref()
    t = R.ref_table_new()
    import ctypes as C
    R.ref_table_update(t, 1, 5, 1)
    R.ref_table_update(t, 1, 3, 4)
    R.ref_table_update(t, 2, 5, 2)
    R.ref_table_update(t, 1, 5, 3)
    assert R.ref_table_lookup(t, 1, 5) == 3
    out = (C.c_int64 * 8)()
    n = R.ref_table_from(t, 1, out, 8)
    assert list(out[:n]) == [5, 3]
    n = R.ref_table_to(t, 5, out, 8)
    assert list(out[:n]) == [1, 2]
    assert R.ref_table_count_from(t, 1) == 2 and R.ref_table_count_to(t, 5) == 2
    # Removal (update with `no`) is not exercised on the reference build: it
    # erases inside its equal_range loop (rts/dependency.cpp:26-33) and
    # crashes; the library's dependency_table erases once, after the lookup.
    R.ref_table_free(t)


def test_dependency_truth_table(pp):
    base = (1 << 20, 64, 64, 4)
    other = (1 << 30, 64, 64, 4)
    # (from read_only, to read_only) -> expected mask on an overlapping pair
    table = {(False, True): 1, (True, False): 2, (False, False): 4, (True, True): 0}
    for (ro_f, ro_t), want in table.items():
        assert pp.check_dependency([(ro_f, base)], [(ro_t, base)]) == want
        assert pp.check_dependency([(ro_f, base)], [(ro_t, other)]) == 0
    # masks OR over all overlapping operand pairs
    got = pp.check_dependency([(False, base), (True, other)], [(True, base), (False, other)])
    assert got == 1 | 2


def _lists(n, edges):
    succ = [[] for _ in range(n)]
    pred = [[] for _ in range(n)]
    for a, b, m in sorted(edges):
        succ[a].append((b, m))
    for a, b, m in sorted(edges, key=lambda e: (e[1], e[0])):
        pred[b].append((a, m))
    return succ, pred


def test_python_scheduler_hand_traced():
    """Schedules derived by hand from rts/scheduler.cpp (see comments in
    oracle/scheduler.py for the rules exercised)."""
    # 0->2, 1->2, 2->3, 0->3 ; critical-node pick of i0, wait insertion, sync elimination
    succ, pred = _lists(4, [(0, 2, 1), (1, 2, 1), (2, 3, 1), (0, 3, 1)])
    out = sref.schedule(4, [100, 200, 300, 50], succ, pred, 2, 10)
    assert out == [[(0, True, 0)], [(1, True, 0), (0, False, 200), (2, True, 200), (3, True, 500)]]
    # 0->1,2,3 ; 1,2->4 ; FLOP floor 20, ready-list prepend order, after-sync worker fallback
    succ, pred = _lists(5, [(0, 1, 1), (0, 2, 1), (0, 3, 1), (1, 4, 1), (2, 4, 1)])
    out = sref.schedule(5, [10] * 5, succ, pred, 2, 20)
    assert sref.text(out) == ["i0 i3 i1 w2 i4", "w0 i2"]
    assert out[0] == [(0, True, 0), (3, True, 20), (1, True, 40), (2, False, 60), (4, True, 60)]
    assert out[1] == [(0, False, 0), (2, True, 20)]


def test_cpp_scheduler_matches_hand_traced(pp):
    s = pp.Schedule([100, 200, 300, 50], [(0, 2), (1, 2), (2, 3), (0, 3)], [1, 1, 1, 1], 2)
    # the library floors with its own _MIN_INST_FLOPS, larger than these costs
    mf = pp.min_inst_flops()
    succ, pred = _lists(4, [(0, 2, 1), (1, 2, 1), (2, 3, 1), (0, 3, 1)])
    want = sref.schedule(4, [100, 200, 300, 50], succ, pred, 2, mf)
    assert [s.worker(w) for w in range(2)] == want


def random_dag(rng, n, p):
    edges = []
    for j in range(n):
        for i in range(j):
            if rng.random() < p:
                edges.append((i, j, int(rng.integers(1, 8))))
    return edges


@pytest.mark.parametrize("seed", range(12))
def test_cpp_scheduler_bit_exact_random(pp, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 120))
    edges = random_dag(rng, n, float(rng.uniform(0.01, 0.3)))
    flops = rng.integers(1, 1 << 22, n)
    if seed % 3 == 0:
        flops[:] = rng.integers(1, 4, n) * (1 << 20)  # many ties
    workers = int(rng.integers(1, 9))
    s = pp.Schedule(flops, [(a, b) for a, b, _ in edges], [m for _, _, m in edges], workers)
    succ, pred = _lists(n, edges)
    want = sref.schedule(n, [int(f) for f in flops], succ, pred, workers, pp.min_inst_flops())
    assert [s.worker(w) for w in range(workers)] == want


def test_schedule_is_valid(pp):
    """Every instruction runs once; every cross-worker predecessor is either
    waited for or already covered by an earlier wait on the same worker."""
    rng = np.random.default_rng(99)
    n = 200
    edges = random_dag(rng, n, 0.05)
    workers = 6
    s = pp.Schedule(rng.integers(1, 1 << 21, n), [(a, b) for a, b, _ in edges], [m for _, _, m in edges], workers)
    where, pos = {}, {}
    for w in range(workers):
        for k, (i, ex, _) in enumerate(s.worker(w)):
            if ex:
                assert i not in where
                where[i], pos[i] = w, k
    assert sorted(where) == list(range(n))
    # simulate: each worker proceeds in order; waits block until done
    done = set()
    ptr = [0] * workers
    streams = [s.worker(w) for w in range(workers)]
    progress = True
    while progress:
        progress = False
        for w in range(workers):
            while ptr[w] < len(streams[w]):
                i, ex, _ = streams[w][ptr[w]]
                if not ex:
                    if i not in done:
                        break
                else:
                    preds = [a for a, b, m in edges if b == i]
                    assert all(p in done for p in preds), f"i{i} ran before a predecessor"
                    done.add(i)
                ptr[w] += 1
                progress = True
    assert len(done) == n
