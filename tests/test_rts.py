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


@pytest.mark.skipif(not __import__("oracle").ref_available(), reason="reference rts library not built here")
@pytest.mark.parametrize("seed", range(6))
def test_dependency_table_matches_reference(pp, orc, seed):
    """The library's dependency_table and the reference's own compiled
    rts/dependency.cpp agree on random insert / retype sequences: lookups,
    successor and predecessor lists (insertion order) and counts
    (rts/dependency.cpp:11-65).  Removals and clear() are left out here: the
    reference's removal erases inside its equal_range loop (undefined
    behaviour) and its clear() keeps the masks; both are checked against a
    model below."""
    import ctypes as C
    R = orc.ref()
    rng = np.random.default_rng(seed)
    ref = R.ref_table_new()
    mine = pp.DepTable()
    n = int(rng.integers(4, 40))
    for _ in range(int(rng.integers(20, 400))):
        f, t, m = int(rng.integers(0, n)), int(rng.integers(0, n)), int(rng.integers(1, 8))
        R.ref_table_update(ref, f, t, m)
        mine.update(f, t, m)
    buf = (C.c_int64 * 4096)()
    for i in range(n):
        k = R.ref_table_from(ref, i, buf, 4096)
        assert mine.dep_from(i) == list(buf[:k])
        k = R.ref_table_to(ref, i, buf, 4096)
        assert mine.dep_to(i) == list(buf[:k])
        assert mine.count_from(i) == R.ref_table_count_from(ref, i)
        assert mine.count_to(i) == R.ref_table_count_to(ref, i)
        for j in range(n):
            assert mine.lookup(i, j) == R.ref_table_lookup(ref, i, j)
    R.ref_table_free(ref)


def test_dependency_table_removal_and_clear(pp):
    """update(no) removes exactly one edge from the edge set and both lists,
    keeping the order of the others; clear() empties every container, so a
    re-inserted edge carries its new mask."""
    rng = np.random.default_rng(5)
    t = pp.DepTable()
    model = {}  # (f, t) -> mask, insertion ordered
    for _ in range(600):
        f, to = int(rng.integers(0, 12)), int(rng.integers(0, 12))
        m = int(rng.integers(0, 8)) if rng.random() < 0.7 else 0
        t.update(f, to, m)
        if m == 0:
            model.pop((f, to), None)
        elif (f, to) in model:
            model[(f, to)] = m
        else:
            model[(f, to)] = m
    for i in range(12):
        assert t.dep_from(i) == [b for (a, b) in model if a == i]
        assert t.dep_to(i) == [a for (a, b) in model if b == i]
        assert t.count_from(i) == sum(a == i for (a, b) in model)
        for j in range(12):
            assert t.lookup(i, j) == model.get((i, j), 0)
    t.clear()
    assert all(t.count_from(i) == 0 and t.count_to(i) == 0 for i in range(12))
    t.update(3, 4, 2)
    assert t.lookup(3, 4) == 2 and t.dep_from(3) == [4]


# Schedules traced by hand through rts/scheduler.cpp, FLOPs in units of F =
# _MIN_INST_FLOPS, two workers.  Each entry is (instruction, execute?, the
# worker's FLOP clock at the entry).
HAND_TRACED = {
    # 0 -> 1, 2, 3; 4 alone.  Step 1 takes i0 over the head i4 (more
    # successors, rts/scheduler.cpp:173-179); its three successors are
    # prepended into one fresh list (equal max_sync, :343-344) which hookup
    # merges behind the head (:142-161); i2 finds no worker with load <= its
    # sync and waits for i0 on worker 1; i1 falls back to worker 0 (the
    # initial earliest_after_sync, :270-286).
    "fanout": (5, [1, 1, 1, 1, 1], [(0, 1), (0, 2), (0, 3)],
               [[(0, True, 0), (3, True, 1), (1, True, 2)],
                [(4, True, 0), (0, False, 1), (2, True, 1)]], 1),
    # 0 (3F) -> 3, 5, 6; 1 -> 2, 4; 2 -> 5; 4 -> 6.  Step 1 takes i0 (3
    # successors) over the head i1; i4 and i2 become ready behind the head
    # i3 although their max_sync is smaller (hookup only inserts after the
    # head); step 3 takes the head i3 because its max_sync exceeds the
    # smallest load (:165-171); step 4 keeps the head i4 on a num_deps tie
    # (strict >); i2 goes to the earliest worker after its sync; i6 and i5
    # each wait once.
    "head": (7, [3, 1, 1, 1, 1, 1, 1], [(0, 3), (0, 5), (0, 6), (1, 2), (1, 4), (2, 5), (4, 6)],
             [[(0, True, 0), (3, True, 3), (2, False, 4), (5, True, 4)],
              [(1, True, 0), (4, True, 1), (2, True, 2), (0, False, 3), (6, True, 3)]], 2),
    # 0 (2F) -> 3, 4; 1 -> 2, 5; 2 -> 3; 3 -> 4.  i5 is prepended before i2;
    # steps 3 and 4 take the second node (more successors) instead of the
    # head i5; i3 waits for i0 (worker 1, clock 2F); i5 runs on the worker
    # after its sync and waits for i1; i4's sync with i0 is dropped because
    # worker 0 already waited for worker 1 at clock 2F (eliminateExtraSync,
    # :289-304).
    "elim": (6, [2, 1, 1, 1, 1, 1], [(0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (3, 4)],
             [[(1, True, 0), (2, True, 1), (0, False, 2), (3, True, 2), (4, True, 3)],
              [(0, True, 0), (1, False, 2), (5, True, 2)]], 2),
}


@pytest.mark.parametrize("name", sorted(HAND_TRACED))
def test_schedules_traced_by_hand(pp, name):
    """The Python restatement and the library's scheduler both reproduce
    schedules derived by hand from rts/scheduler.cpp, including hookup of
    several fresh entries, num_deps ties, the head-only insertion and
    eliminated syncs."""
    n, units, edges, want, waits = HAND_TRACED[name]
    F = pp.min_inst_flops()
    flops = [u * F for u in units]
    want = [[(i, ex, c * F) for i, ex, c in w] for w in want]
    succ, pred = _lists(n, [(a, b, 1) for a, b in edges])
    assert sref.schedule(n, flops, succ, pred, 2, F) == want
    s = pp.Schedule(flops, edges, [1] * len(edges), 2)
    assert [s.worker(w) for w in range(2)] == want
    assert s.comm_waits() == waits
