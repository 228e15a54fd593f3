"""Serving loop admission logic on the CPU (device calls replaced by fakes).

Failure handling exercised here, never on the GPU:
  * an HBM budget too small for two matrices at once: requests queue and
    idle resident factorizations are evicted (LRU), nothing over-allocates;
  * a request whose solver construction fails is reported and its
    reservation released, and the server keeps serving;
  * reuse requests never refactor.
"""
import importlib
import threading
import time

import numpy as np

serve = importlib.import_module("pastel-palettes_amd.serve")


class FakeHM:
    def __init__(self, n):
        self.n = n
        self.storage = n * 40


class FakeCtx:
    def __init__(self):
        self.used = 0
        self.peak = 0
        self.lock = threading.Lock()

    def max_workers(self, precision, wavefronts):
        return 512

    def bytes_in_use(self):
        return self.used

    def alloc(self, b):
        with self.lock:
            self.used += b
            self.peak = max(self.peak, self.used)

    def release(self, b):
        with self.lock:
            self.used -= b


class FakeSolver:
    fail_keys = set()
    fail_factor_keys = set()
    factor_calls = 0
    closed = 0
    size_fn = staticmethod(lambda hm: hm.storage * 12)

    def __init__(self, ctx, hm, ptr, precision, workers, wavefronts, nrhs_max):
        if hm.n in FakeSolver.fail_keys:
            raise RuntimeError("construction failed")
        self.ctx, self.hm = ctx, hm
        self.bytes = FakeSolver.size_fn(hm)
        ctx.alloc(self.bytes)

    def factor(self, stream=None):
        FakeSolver.factor_calls += 1
        time.sleep(0.002)
        if self.hm.n in FakeSolver.fail_factor_keys:
            raise RuntimeError("VM watchdog")

    def solve(self, b_ptr, x_ptr, nrhs, tol, max_iter, hip_stream=None):
        time.sleep(0.001)
        return {"iterations": 1, "converged": 1, "residual": 1e-12}

    def close(self):
        FakeSolver.closed += 1
        self.ctx.release(self.bytes)


class FakeTensor:
    def __init__(self, a):
        self.a = np.asarray(a)

    def to(self, *a, **k):
        return self

    def data_ptr(self):
        return 0

    def cpu(self):
        return self

    def numpy(self):
        return self.a


class FakeTorch:
    class cuda:
        class Stream:
            cuda_stream = None

            def synchronize(self):
                pass

        class _Ctx:
            def __enter__(self):
                return self

            def __exit__(self, *a):
                return False

        @staticmethod
        def stream(s):
            return FakeTorch.cuda._Ctx()

    @staticmethod
    def from_numpy(a):
        return FakeTensor(a)

    @staticmethod
    def zeros_like(t):
        return FakeTensor(np.zeros_like(t.a))


class FakePP:
    FP64, MIXED = 0, 2
    Solver = FakeSolver


class Arena:
    def data_ptr(self):
        return 0


def make_server(budget, slots=2, sizes=(1000, 2000, 3000)):
    ctx = FakeCtx()
    mats = {f"m{n}": (FakeHM(n), Arena()) for n in sizes}
    srv = serve.Server(FakePP, FakeTorch, ctx, mats, FakePP.MIXED, budget_bytes=budget, slots=slots,
                       workers_per_slot=8, nrhs_max=2)
    return srv, ctx, mats


def test_budget_forces_eviction_not_overcommit():
    FakeSolver.fail_keys = set()
    srv, ctx, mats = make_server(budget=0)
    one = max(srv._estimate(k) for k in mats)
    srv.budget = int(1.5 * one)  # room for one resident matrix at a time
    FakeSolver.size_fn = staticmethod(lambda hm: srv._estimate(f"m{hm.n}"))
    reqs = [serve.Request(i, key, np.ones((5, 1))) for i, key in enumerate(["m1000", "m2000", "m3000", "m1000"])]
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        assert r.event.wait(10)
    srv.close()
    assert all(r.error is None for r in reqs)
    assert ctx.peak <= srv.budget
    assert srv.evictions >= 2
    FakeSolver.size_fn = staticmethod(lambda hm: hm.storage * 12)


def test_failed_construction_is_reported_and_released():
    FakeSolver.fail_keys = {2000}
    srv, ctx, mats = make_server(budget=1 << 40)
    reqs = [serve.Request(0, "m2000", np.ones((5, 1))), serve.Request(1, "m1000", np.ones((5, 1)))]
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        assert r.event.wait(10)
    assert reqs[0].error is not None and reqs[1].error is None
    assert srv.reserved == 0 and "m2000" not in srv.resident
    srv.close()
    FakeSolver.fail_keys = set()


def test_reuse_does_not_refactor():
    srv, ctx, mats = make_server(budget=1 << 40, slots=1)
    FakeSolver.factor_calls = 0
    first = serve.Request(0, "m1000", np.ones((5, 1)))
    srv.submit(first)
    first.event.wait(10)
    again = [serve.Request(i, "m1000", np.ones((5, 1)), reuse=True) for i in range(1, 4)]
    for r in again:
        srv.submit(r)
    for r in again:
        r.event.wait(10)
    srv.close()
    assert FakeSolver.factor_calls == 1
    assert all(r.factored is False for r in again)


def test_capacity_check():
    ctx = FakeCtx()
    try:
        serve.Server(FakePP, FakeTorch, ctx, {}, 0, 1 << 30, slots=8, workers_per_slot=128)
    except ValueError:
        return
    raise AssertionError("slots x workers above co-resident capacity must be refused")


def test_poisson_trace_shape():
    tr = serve.poisson_trace(200.0, 2.0, ["a", "b", "c"], seed=1)
    assert 300 < len(tr) < 500
    assert all(1 <= t[2] <= 64 for t in tr)
    frac = sum(t[3] for t in tr) / len(tr)
    assert 0.2 < frac < 0.4


def test_failed_factorization_evicts_the_entry():
    """A factorization that raises after its solver was built leaves nothing
    resident: the solver is closed and its bytes released, and a later reuse
    request for the same key builds and factors again instead of solving
    with invalid factors."""
    FakeSolver.fail_keys = set()
    FakeSolver.fail_factor_keys = {1000}
    FakeSolver.factor_calls = 0
    FakeSolver.closed = 0
    srv, ctx, mats = make_server(budget=1 << 40, slots=1)
    bad = serve.Request(0, "m1000", np.ones((5, 1)))
    srv.submit(bad)
    assert bad.event.wait(10)
    assert bad.error is not None and "m1000" not in srv.resident
    assert ctx.used == 0 and FakeSolver.closed == 1 and srv.bytes_resident == 0
    FakeSolver.fail_factor_keys = set()
    again = serve.Request(1, "m1000", np.ones((5, 1)), reuse=True)
    srv.submit(again)
    assert again.event.wait(10)
    srv.close()
    assert again.error is None and again.factored
    assert FakeSolver.factor_calls == 2


class TimedSolver(FakeSolver):
    """Fake solver whose factorization takes time proportional to n and
    records how many workers are on the device at once."""
    active = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, ctx, hm, ptr, precision, workers, wavefronts, nrhs_max):
        super().__init__(ctx, hm, ptr, precision, workers, wavefronts, nrhs_max)
        self.workers = workers

    def factor(self, stream=None):
        with TimedSolver.lock:
            TimedSolver.active += self.workers
            TimedSolver.peak = max(TimedSolver.peak, TimedSolver.active)
        time.sleep(self.hm.n * 1e-6)
        with TimedSolver.lock:
            TimedSolver.active -= self.workers


class TimedPP(FakePP):
    Solver = TimedSolver


def _timed_server(policy, slots, share, sizes):
    ctx = FakeCtx()
    mats = {f"m{n}": (FakeHM(n), Arena()) for n in sizes}
    srv = serve.Server(TimedPP, FakeTorch, ctx, mats, FakePP.MIXED, budget_bytes=1 << 40, slots=slots,
                       workers_per_slot=32, nrhs_max=2, share=share, policy=policy)
    return srv


def test_worker_pool_never_overcommitted():
    """With per-request worker shares the workers of the factorizations on
    the device never exceed the pool (slots x workers per slot), so every
    persistent launch is co-resident."""
    TimedSolver.active = TimedSolver.peak = 0
    sizes = (1000, 4000, 8000, 16000, 32000)
    share = lambda n: 8 if n < 5000 else 64  # noqa: E731
    srv = _timed_server("sjf", 4, share, sizes)
    assert srv.pool == 128
    reqs = [serve.Request(i, f"m{sizes[i % len(sizes)]}", np.ones((5, 1))) for i in range(20)]
    for r in reqs:
        srv.submit(r)
    for r in reqs:
        assert r.event.wait(20)
    srv.close()
    assert all(r.error is None for r in reqs)
    assert 0 < TimedSolver.peak <= 128
    assert srv.free_workers == srv.pool


def test_sjf_serves_short_requests_first():
    """One slot, everything queued at once: shortest job first finishes the
    small factorizations before the large one that arrived with them, and
    FIFO keeps arrival order."""
    sizes = (1000, 20000)
    order = {}
    for policy in ("fifo", "sjf"):
        srv = _timed_server(policy, 1, None, sizes)
        srv.lock.acquire()  # queue everything before the slot takes any
        reqs = [serve.Request(0, "m20000", np.ones((5, 1)))]
        reqs += [serve.Request(i, "m1000", np.ones((5, 1))) for i in range(1, 5)]
        for r in reqs:
            srv.queue.append(r)
        srv.lock.release()
        with srv.lock:
            srv.lock.notify_all()
        for r in reqs:
            assert r.event.wait(20)
        srv.close()
        order[policy] = [r.rid for r in sorted(reqs, key=lambda r: r.done)]
    assert order["fifo"][0] == 0
    assert order["sjf"][-1] == 0


def test_sjf_aging_serves_a_waiting_long_request():
    """A long request is not starved by a steady stream of short ones: its
    priority grows with its time in the queue."""
    srv = _timed_server("sjf", 1, None, (1000, 20000))
    srv.aging = 50.0
    long_req = serve.Request(0, "m20000", np.ones((5, 1)))
    srv.submit(long_req)
    shorts = []
    t0 = time.perf_counter()
    while not long_req.event.is_set() and time.perf_counter() - t0 < 5.0:
        r = serve.Request(len(shorts) + 1, "m1000", np.ones((5, 1)))
        shorts.append(srv.submit(r))
        time.sleep(0.0005)
    assert long_req.event.wait(10)
    for r in shorts:
        r.event.wait(10)
    srv.close()
    assert long_req.error is None
