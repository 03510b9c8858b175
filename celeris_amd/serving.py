"""Serving loop: continuous batching of scenario requests on one GPU.

Requests (a scenario and a number of steps) arrive over time.  The server keeps every
admitted scenario resident in one Batch and advances them together in blocks of steps;
arrived requests are admitted while their device footprint fits the HBM budget (first come,
first served; least work first under "srpt") and finished ones retire, so requests that
arrive while others run join the running batch.  Admission overlaps the GPU: a block is
launched (asynchronously), then at most one waiting request is set up on the host (its
engine, tables and uploads) while that block runs, then the loop waits for the block.  A
request's latency is its completion time minus its arrival time; a member whose state turns
non-finite (the kernels' numeric sentinel) is evicted and recorded as failed, the others
continue.

Two policies decide which resident requests advance in a block.  "fair" advances all of
them (processor sharing).  "srpt" advances the requests with the least remaining work
(cells x remaining steps) first, taking them in that order until their cells fill one
large grid's worth of the GPU (gpu_cells); the rest stay resident but paused.

The comparison point is the plain engine serving the same stream one request at a time
(first come, first served): each request's time alone, from creating its engine to having
stepped it and released it, is measured on the GPU, and the queue is replayed from those
times.  Both sides pay the same per-request setup (engine creation, uploads, tables).
"""
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import Batch, ShallowWaterEngine, estimate_device_bytes

HBM_BYTES = 288 * 1024 ** 3       # MI355X HBM3E capacity


@dataclass
class Record:
    rid: int
    cells: int
    steps: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    failed: bool = False

    @property
    def latency(self):
        return self.finished - self.arrival


def summarize(records, makespan):
    done = [r for r in records if not r.failed]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    work = sum(r.cells * r.steps for r in done)
    return {"requests": len(done), "failed": len(records) - len(done),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "makespan_s": makespan,
            "gcell_steps_per_s": work / makespan / 1e9, "cell_steps": work}


def serve(requests, hbm_budget=HBM_BYTES, block=50, max_active=512, keep_states=False, graph_steps=0,
          policy="fair", gpu_cells=2048 * 2048, clock=time.perf_counter, sleep=time.sleep):
    """Runs the request stream in real time through one Batch.  Each iteration launches a
    block that ends when the first running request completes, at most `block` steps long
    while requests are still to arrive or wait for admission (so that they join soon), sets
    up at most one waiting request while it runs, then waits for it.  Every block ends in a
    synchronisation and a new work-list, so blocks are no shorter than they need to be.  graph_steps > 0 replays
    captured blocks of that many steps (the setup then waits for the block).  policy "fair"
    or "srpt" (see the module notes).  Returns (records, summary, states) where states maps
    request id to its final state (FP64 view of the state, if keep_states)."""
    if policy not in ("fair", "srpt"):
        raise ValueError("policy must be 'fair' or 'srpt'")
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    queue = deque()
    active = {}                    # batch id -> [request, remaining steps, bytes, record]
    records = []
    states = {}
    used = 0
    b = Batch(graph_steps=graph_steps)
    t0 = clock()

    def arrivals():
        now = clock() - t0
        while pending and pending[0].arrival <= now:
            queue.append(pending.popleft())

    def admit_one():
        """Sets up the next waiting request if its footprint fits: the first to arrive
        ("fair") or the one with the least work ("srpt"); returns True if it did."""
        nonlocal used
        if not queue or len(active) >= max_active:
            return False
        req = queue[0]
        if policy == "srpt":
            req = min(queue, key=lambda r: r.scenario.s.nx * r.scenario.s.ny * r.steps)
        need = estimate_device_bytes(req.scenario.s, req.scenario.boundaries)
        if used + need > hbm_budget and active:
            return False
        queue.remove(req)
        sc = req.scenario
        sid = b.add(sc.s, sc.boundaries, sc.bottom, sc.state)
        rec = Record(req.rid, sc.s.nx * sc.s.ny, req.steps, req.arrival, admitted=clock() - t0)
        active[sid] = [req, req.steps, need, rec]
        used += need
        return True

    def retire(sid, at, failed=False):
        nonlocal used
        v = active.pop(sid)
        v[3].finished = at
        v[3].failed = failed
        if keep_states and not failed:
            states[v[0].rid] = b.get_state_f64(sid)
        b.remove(sid)
        used -= v[2]
        records.append(v[3])

    # a captured block fails if an engine is set up while it is captured, so with
    # graph_steps the setup waits for the block
    overlap = graph_steps == 0
    try:
        while pending or queue or active:
            arrivals()
            if not active:
                if not admit_one():
                    if pending:
                        sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                    continue
            run = list(active)
            if policy == "srpt":
                run, cells = [], 0
                for sid in sorted(active, key=lambda i: active[i][3].cells * active[i][1]):
                    if run and cells + active[sid][3].cells > gpu_cells:
                        break
                    run.append(sid)
                    cells += active[sid][3].cells
                for sid in active:
                    b.set_paused(sid, sid not in run)
            k = min(active[sid][1] for sid in run)
            if pending or queue:
                k = min(k, block)
            b.step(k)                     # asynchronous: the block runs while the host sets up
            if overlap:
                arrivals()
                admit_one()
            b.sync()
            if not overlap:
                arrivals()
                admit_one()
            done_at = clock() - t0
            failed = set(b.failed())
            for sid in run:
                active[sid][1] -= k
            for sid in list(active):
                if sid in failed:
                    retire(sid, done_at, failed=True)
                elif sid in run and active[sid][1] == 0:
                    retire(sid, done_at)
    finally:
        b.close()
    makespan = max(r.finished for r in records) - min(r.arrival for r in records)
    return records, summarize(records, makespan), states


def alone_run_time(req, graph_steps=0):
    """Time of one request served alone with the plain engine (seconds): creating its engine
    (allocation, uploads, tables), stepping it and releasing it -- the same per-request work
    the batched server does for it."""
    sc = req.scenario
    t = time.perf_counter()
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    if graph_steps:
        e.set_graph_steps(graph_steps)
    e.timestep(req.steps)
    e.sync()
    e.close()
    return time.perf_counter() - t


def one_at_a_time(requests, graph_steps=0, run_times=None):
    """FIFO replay of the stream on the plain engine: each request's measured time alone
    (or run_times[rid] if given, measured the same way), queued behind the previous one."""
    recs = []
    free_at = 0.0
    for req in sorted(requests, key=lambda r: r.arrival):
        run = run_times[req.rid] if run_times is not None else alone_run_time(req, graph_steps)
        start = max(req.arrival, free_at)
        free_at = start + run
        sc = req.scenario
        recs.append(Record(req.rid, sc.s.nx * sc.s.ny, req.steps, req.arrival, admitted=start, finished=free_at))
    makespan = max(r.finished for r in recs) - min(r.arrival for r in recs)
    return recs, summarize(recs, makespan)


def throughput_bound(requests, run_times, best_rate):
    """Upper bound of the batched/one-at-a-time throughput ratio for a saturated stream:
    one-at-a-time takes the sum of the measured times alone, and no batch can advance the
    requests' cell-steps faster than best_rate (cell-steps/s, the best rate measured for any
    single grid), so the ratio is at most sum(run_times) / (total cell-steps / best_rate)."""
    work = sum(r.scenario.s.nx * r.scenario.s.ny * r.steps for r in requests)
    return sum(run_times[r.rid] for r in requests) / (work / best_rate)
