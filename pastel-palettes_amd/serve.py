"""Serving loop: continuous admission of factor + solve requests.

A request names an H-matrix (resident FP64 arena on the device), carries
1..64 right-hand sides, and either reuses a resident factorization of that
matrix or asks for a fresh one.  The server keeps `slots` HIP streams; each
slot runs one request at a time on its own stream with `workers_per_slot`
persistent workers, so requests that arrive while others are on the GPU are
admitted as soon as a slot is free and their work overlaps on the device.
Admission is by an HBM budget: a request whose solver does not fit waits,
after resident factorizations that are not in use have been evicted (least
recently used first).  Factorizations stay resident for later requests that
reuse them.

The slot count times the workers per slot never exceeds the number of VM
workgroups that fit co-resident on the device (the persistent kernels of
all slots must be resident together).  That product is the server's worker
pool.  With a `share` function each request's factorization gets a worker
count sized by its matrix instead of a whole slot (a small matrix's DAG has
little parallelism to spread), and a request is admitted only while the
workers of the requests on the device, its own included, fit in the pool,
so their persistent launches are always co-resident.

Admission order: "fifo" takes the first admissible request in arrival
order; "sjf" takes the admissible request with the smallest estimated
service time less `aging` times its time in the queue (shortest job first,
and a long job's priority grows as it waits, so it is not starved).  The
estimate of a key's factor + solve and of its solve alone starts from a size
model and follows the measured service times (exponential average).

Every request's arithmetic is independent of the others (static schedule,
one workgroup per instruction, fixed reduction orders), so its result is the
same whether it runs alone or in the mix.
"""
import collections
import heapq
import threading
import time

import numpy as np


class Request:
    __slots__ = ("rid", "key", "nrhs", "reuse", "arrival", "b", "x", "start", "done", "report", "error",
                 "factored", "event", "slot")

    def __init__(self, rid, key, b, reuse=False, arrival=None):
        self.rid, self.key, self.b, self.reuse = rid, key, b, reuse
        self.nrhs = b.shape[1]
        self.arrival = time.perf_counter() if arrival is None else arrival
        self.start = self.done = None
        self.report = self.error = None
        self.factored = False
        self.x = None
        self.event = threading.Event()
        self.slot = -1

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Resident:
    __slots__ = ("solver", "bytes", "in_use", "last_used", "key", "workers")

    def __init__(self, key, solver, nbytes, workers):
        self.key, self.solver, self.bytes, self.workers = key, solver, nbytes, workers
        self.in_use = 0
        self.last_used = time.perf_counter()


def default_share(n, leaf=256):
    """Workers for a factorization of order n: four per diagonal leaf, at
    least 16 and at most 128 (the DAG of a HODLR factorization has about as
    many independent low-rank updates per step as leaves below the current
    one)."""
    return int(min(128, max(16, 4 * n // leaf)))


class Server:
    """HBM-budgeted, multi-slot request loop over pp.Solver."""

    def __init__(self, pp, torch, ctx, matrices, precision, budget_bytes, slots=4, workers_per_slot=64,
                 wavefronts=4, tol=1e-10, max_iter=60, nrhs_max=64, share=None, policy="fifo", aging=1.0):
        cap = ctx.max_workers(precision, wavefronts)
        if slots * workers_per_slot > cap:
            raise ValueError(f"{slots} slots x {workers_per_slot} workers exceed co-resident capacity {cap}")
        if policy not in ("fifo", "sjf"):
            raise ValueError(f"unknown admission policy {policy!r}")
        self.pool = slots * workers_per_slot
        self.free_workers = self.pool
        self.share = share
        self.policy, self.aging = policy, aging
        self.est = {}  # (key, reuse) -> exponential average of the service time (s)
        self.pp, self.torch, self.ctx = pp, torch, ctx
        self.matrices = matrices  # key -> (hm, device FP64 arena tensor)
        self.precision, self.budget = precision, budget_bytes
        self.workers, self.wavefronts, self.tol, self.max_iter = workers_per_slot, wavefronts, tol, max_iter
        self.nrhs_max = nrhs_max
        self.lock = threading.Condition()
        self.queue = collections.deque()
        self.resident = {}  # key -> _Resident
        self.bytes_resident = 0
        self.reserved = 0  # bytes admitted but not yet allocated
        self.stop = False
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        self.threads = [threading.Thread(target=self._run, args=(i,), daemon=True) for i in range(slots)]
        self.completed = []
        self.evictions = 0
        for t in self.threads:
            t.start()

    # ------------------------------------------------------------ admission
    def submit(self, req):
        with self.lock:
            self.queue.append(req)
            self.lock.notify_all()
        return req

    def close(self):
        with self.lock:
            self.stop = True
            self.lock.notify_all()
        for t in self.threads:
            t.join()
        for r in list(self.resident.values()):
            r.solver.close()
        self.resident.clear()

    def _estimate(self, key, workers=None):
        hm, _ = self.matrices[key]
        workers = self.workers if workers is None else workers
        eb = 8 if self.precision == self.pp.FP64 else 4
        n = hm.n
        # FP64 arena + factor arena (both with the leaf-inverse area) + rhs/residual
        # buffers + per-worker scratch (upper bound)
        elems = getattr(hm, "solver_elems", hm.storage)
        return elems * (8 + eb) + n * self.nrhs_max * (eb + 8) + workers * 64 * 1024 * 1024 // 16

    def _evict_for(self, need):
        """Evict idle resident factorizations (LRU) until `need` bytes fit."""
        fits = lambda: self.ctx.bytes_in_use() + self.reserved + need <= self.budget
        idle = sorted((r for r in self.resident.values() if r.in_use == 0), key=lambda r: r.last_used)
        for r in idle:
            if fits():
                break
            r.solver.close()
            del self.resident[r.key]
            self.bytes_resident -= r.bytes
            self.evictions += 1
        return fits()

    def _workers_for(self, key):
        if self.share is None:
            return self.workers
        return int(min(self.pool, max(1, self.share(self.matrices[key][0].n))))

    def _service_estimate(self, req):
        """Estimated service time (s) of a request: measured average of its
        key (factor + solve, or solve alone for a resident factorization),
        else a size model (a chain of diagonal leaves for the factorization)."""
        res = self.resident.get(req.key)
        solve_only = req.reuse and res is not None and res.solver is not None
        k = (req.key, solve_only)
        if k in self.est:
            return self.est[k]
        n = self.matrices[req.key][0].n
        return 2e-3 + (0.0 if solve_only else 0.35e-3 * n / 256.0)

    def _order(self):
        """Queue positions in the order the policy offers them."""
        if self.policy == "fifo":
            return range(len(self.queue))
        now = time.perf_counter()
        pri = [self._service_estimate(q) - self.aging * (now - q.arrival) for q in self.queue]
        return sorted(range(len(self.queue)), key=lambda i: pri[i])

    def _take(self):
        """Next admissible request and its resident entry (or None to build)."""
        while True:
            if self.stop and not self.queue:
                return None, None
            for i in self._order():
                req = self.queue[i]
                res = self.resident.get(req.key)
                if res is not None and res.in_use == 0 and res.solver is not None:
                    if res.workers > self.free_workers:
                        break  # the first in order waits for workers; no one passes it
                    del self.queue[i]
                    res.in_use += 1
                    self.free_workers -= res.workers
                    return req, res
                if res is None:
                    w = self._workers_for(req.key)
                    if w > self.free_workers:
                        break
                    need = self._estimate(req.key, w)
                    if self._evict_for(need):
                        del self.queue[i]
                        # reserve the key so no other slot builds it concurrently
                        placeholder = _Resident(req.key, None, need, w)
                        placeholder.in_use = 1
                        self.resident[req.key] = placeholder
                        self.bytes_resident += need
                        self.reserved += need
                        self.free_workers -= w
                        return req, placeholder
            self.lock.wait(timeout=0.05)

    # ------------------------------------------------------------ execution
    def _run(self, slot):
        torch, pp = self.torch, self.pp
        st = self.streams[slot]
        while True:
            with self.lock:
                req, res = self._take()
            if req is None:
                return
            req.slot = slot
            req.start = time.perf_counter()
            try:
                if res.solver is None:
                    hm, arena = self.matrices[req.key]
                    try:
                        res.solver = pp.Solver(self.ctx, hm, arena.data_ptr(), precision=self.precision,
                                               workers=res.workers, wavefronts=self.wavefronts,
                                               nrhs_max=self.nrhs_max)
                    finally:
                        with self.lock:
                            self.reserved -= res.bytes
                    res.solver.factor(st.cuda_stream)
                    req.factored = True
                elif not req.reuse:
                    res.solver.factor(st.cuda_stream)
                    req.factored = True
                with torch.cuda.stream(st):
                    db = torch.from_numpy(req.b).to("cuda", non_blocking=False)
                    dx = torch.zeros_like(db)
                st.synchronize()
                req.report = res.solver.solve(db.data_ptr(), dx.data_ptr(), req.nrhs, tol=self.tol,
                                              max_iter=self.max_iter, hip_stream=st.cuda_stream)
                req.x = dx.cpu().numpy()
            except Exception as e:  # reported per request; the server keeps running
                req.error = e
                # the entry is this request's alone (in_use == 1): after any
                # failure its factors cannot be trusted, so it is evicted and
                # a later request for the key builds and factors afresh
                with self.lock:
                    if res.solver is not None:
                        try:
                            res.solver.close()
                        except Exception:
                            pass
                        res.solver = None
                    if self.resident.get(req.key) is res:
                        del self.resident[req.key]
                        self.bytes_resident -= res.bytes
            req.done = time.perf_counter()
            with self.lock:
                res.in_use -= 1
                res.last_used = req.done
                self.free_workers += res.workers
                if req.error is None:
                    k = (req.key, not req.factored)
                    dt = req.done - req.start
                    self.est[k] = dt if k not in self.est else 0.7 * self.est[k] + 0.3 * dt
                self.completed.append(req)
                self.lock.notify_all()
            req.event.set()


def poisson_trace(rate, duration, keys, nrhs_range=(1, 64), reuse_frac=0.3, seed=0):
    """(arrival_offset, key, nrhs, reuse) tuples of a Poisson stream."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    seen = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        reuse = bool(seen) and rng.random() < reuse_frac
        key = seen[int(rng.integers(len(seen)))] if reuse else keys[int(rng.integers(len(keys)))]
        if not reuse:
            seen.append(key)
        nrhs = int(rng.integers(nrhs_range[0], nrhs_range[1] + 1))
        out.append((t, key, nrhs, reuse))
    return out


def replay(server, trace, make_rhs, realtime=True):
    """Submit a trace (arrival offsets in seconds) and wait for completion;
    returns the requests."""
    t0 = time.perf_counter()
    reqs = []
    for k, (off, key, nrhs, reuse) in enumerate(trace):
        if realtime:
            delay = t0 + off - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        arrival = t0 + off if realtime else time.perf_counter()
        reqs.append(server.submit(Request(k, key, make_rhs(key, nrhs, k), reuse=reuse, arrival=arrival)))
    for r in reqs:
        r.event.wait()
    return reqs, time.perf_counter() - t0


def summarize(reqs, wall):
    lat = np.array([r.latency for r in reqs if r.error is None])
    ok = [r for r in reqs if r.error is None]
    return {
        "requests": len(reqs),
        "errors": sum(r.error is not None for r in reqs),
        "throughput_rps": len(ok) / wall if wall > 0 else 0.0,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": 1e3 * float(np.percentile(lat, 99)) if len(lat) else None,
        "max_residual": max((r.report["residual"] for r in ok), default=None),
        "converged": all(r.report["converged"] for r in ok),
        "factored": sum(r.factored for r in ok),
    }
