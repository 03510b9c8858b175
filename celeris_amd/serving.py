"""Serving loop: continuous batching of scenario requests on one GPU.

Requests (a scenario and a number of steps) arrive over time.  The server keeps every
admitted scenario resident in one Batch and advances all of them together in blocks of
steps; between blocks it admits arrived requests while their device footprint fits the
HBM budget (first come, first served) and retires finished ones, so requests that arrive
while others run join the running batch.  A request's latency is its completion time
minus its arrival time.

Two policies decide which resident requests advance in a block.  "fair" advances all of
them (processor sharing).  "srpt" advances the requests with the least remaining work
(cells x remaining steps) first, taking them in that order until their cells fill one
large grid's worth of the GPU (gpu_cells); the rest stay resident but paused.  Small and
nearly finished requests then complete without waiting for large ones, while the batch
still holds enough cells to keep the GPU busy.

The comparison point is the plain engine serving the same stream one request at a time
(first come, first served): each request's run time alone is measured on the GPU and the
queue is replayed from those times.
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

    @property
    def latency(self):
        return self.finished - self.arrival


def summarize(records, makespan):
    lat = np.array([r.latency for r in records])
    work = sum(r.cells * r.steps for r in records)
    return {"requests": len(records), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
            "makespan_s": makespan, "gcell_steps_per_s": work / makespan / 1e9, "cell_steps": work}


def serve(requests, hbm_budget=HBM_BYTES, block=50, max_active=512, keep_states=False, graph_steps=0,
          policy="fair", gpu_cells=2048 * 2048, clock=time.perf_counter, sleep=time.sleep):
    """Runs the request stream in real time through one Batch, synchronising every `block`
    steps (or sooner, when a request completes).  graph_steps > 0 replays captured blocks
    of that many steps (on MI355X direct launches measured faster).  policy "fair" or
    "srpt" (see the module notes).  Returns (records, summary, states) where states maps
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
    try:
        while pending or queue or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # admission: FIFO while the footprint fits the budget
            while queue and len(active) < max_active:
                req = queue[0]
                need = estimate_device_bytes(req.scenario.s, req.scenario.boundaries)
                if used + need > hbm_budget and active:
                    break
                queue.popleft()
                sc = req.scenario
                sid = b.add(sc.s, sc.boundaries, sc.bottom, sc.state)
                rec = Record(req.rid, sc.s.nx * sc.s.ny, req.steps, req.arrival, admitted=clock() - t0)
                active[sid] = [req, req.steps, need, rec]
                used += need
            if not active:
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
            k = min(block, min(active[sid][1] for sid in run))
            b.step(k)
            b.sync()
            done_at = clock() - t0
            for sid in run:
                v = active[sid]
                v[1] -= k
                if v[1] == 0:
                    v[3].finished = done_at
                    if keep_states:
                        states[v[0].rid] = b.get_state_f64(sid)
                    b.remove(sid)
                    used -= v[2]
                    records.append(v[3])
                    del active[sid]
    finally:
        b.close()
    makespan = max(r.finished for r in records) - min(r.arrival for r in records)
    return records, summarize(records, makespan), states


def alone_run_time(req, graph_steps=0):
    """Device time of one request run alone with the plain engine (seconds)."""
    sc = req.scenario
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    if graph_steps:
        e.set_graph_steps(graph_steps)
    t = time.perf_counter()
    e.timestep(req.steps)
    e.sync()
    dt = time.perf_counter() - t
    e.close()
    return dt


def one_at_a_time(requests, graph_steps=0, run_times=None):
    """FIFO replay of the stream on the plain engine: each request's measured run time alone
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
