"""Pure-Python restatement of the reference list scheduler
(rts/scheduler.cpp) -- TEST INFRASTRUCTURE ONLY.

The DAG is given as successor / predecessor lists in ascending order with
their dependency masks; the output is, per worker, the entries
(instruction, execute?, worker FLOP clock at append).
"""


class _Entry:
    __slots__ = ("inst", "num_deps", "flops", "max_sync", "sync_with")


def schedule(n, flops, succ, pred, workers, min_flops):
    """succ[i] / pred[j]: lists of (node, mask) in ascending node order."""
    indeg = [sum(1 for _, m in pred[j] if m > 0) for j in range(n)]
    flops_after = [0] * n
    executed_by = [-1] * n
    load = [0] * workers
    last_sync = [[0] * workers for _ in range(workers)]
    out = [[] for _ in range(workers)]

    def make(i):
        # rts/scheduler.cpp:70-102
        e = _Entry()
        e.inst = i
        e.num_deps = sum(1 for _, m in succ[i] if m > 0)  # getDepCount_From
        e.flops = flops[i]
        best = [0] * workers
        e.sync_with = [-1] * workers
        e.max_sync = 0
        for src, m in pred[i]:
            if m <= 0:
                continue
            w, f = executed_by[src], flops_after[src]
            if best[w] < f:
                best[w] = f
                e.sync_with[w] = src
            if e.max_sync < f:
                e.max_sync = f
        return e

    def hookup(lst, items):
        # rts/scheduler.cpp:142-161: merge after the head only
        pos = 1
        for q in items:
            while pos < len(lst) and lst[pos].max_sync <= q.max_sync:
                pos += 1
            lst.insert(pos, q)
            pos += 1

    ready = []
    for i in range(n):  # roots prepended -> descending order
        if indeg[i] == 0:
            ready.insert(0, make(i))

    for _ in range(n):
        if not ready:
            raise RuntimeError("cycle")
        smallest = min(load)
        # rts/scheduler.cpp:163-200
        head = ready[0]
        pick = 0
        if not (smallest < head.max_sync):
            best = head.num_deps
            for idx in range(1, len(ready)):
                if ready[idx].max_sync > smallest:
                    break
                if ready[idx].num_deps > best:
                    best = ready[idx].num_deps
                    pick = idx
        e = ready.pop(pick)
        f = max(e.flops, min_flops)
        # rts/scheduler.cpp:267-287
        before, after = -1, 0
        before_load, after_load = -1, load[0]
        for w in range(workers):
            if load[w] <= e.max_sync:
                if load[w] > before_load:
                    before, before_load = w, load[w]
            elif load[w] < after_load:
                after, after_load = w, load[w]
        wk = before if before >= 0 else after
        # rts/scheduler.cpp:289-304
        sw = list(e.sync_with)
        sw[wk] = -1
        for w in range(workers):
            if sw[w] >= 0 and flops_after[sw[w]] <= last_sync[wk][w]:
                sw[w] = -1
        for w in range(workers):
            if sw[w] >= 0:  # rts/scheduler.cpp:317-328
                src = sw[w]
                out[wk].append((src, False, load[wk]))
                load[wk] = max(load[wk], flops_after[src])
                last_sync[wk][executed_by[src]] = flops_after[src]
        out[wk].append((e.inst, True, load[wk]))  # rts/scheduler.cpp:306-315
        load[wk] += f
        flops_after[e.inst] = load[wk]
        executed_by[e.inst] = wk
        # rts/scheduler.cpp:330-370
        fresh = []
        for s, m in succ[e.inst]:
            if m <= 0:
                continue
            indeg[s] -= 1
            if indeg[s] == 0:
                x = make(s)
                if not fresh:
                    fresh = [x]
                elif x.max_sync <= fresh[0].max_sync:
                    fresh.insert(0, x)
                else:
                    hookup(fresh, [x])
        if not ready:
            ready = fresh
        elif fresh:
            hookup(ready, fresh)
    return out


def text(out):
    return [" ".join(("i" if ex else "w") + str(i) for i, ex, _ in w) for w in out]
