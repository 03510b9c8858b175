"""Serving loop: continuous batching of a request stream (celeris_amd.serving).

GPU: every request's result is bit-identical alone and in the mix (the "alone vs mixed"
check), also when the HBM budget forces requests to wait for admission; the one-at-a-time
replay is a FIFO queue over measured run times.  CPU: the request generator's
distributions and the latency summary.
"""
import numpy as np
import pytest

from celeris_amd import ShallowWaterEngine, estimate_device_bytes, serving
from celeris_amd import scenarios as S


def small_stream(n=6, rate=200.0, seed=3):
    return S.serving_requests(n, rate, seed=seed, sides=(48, 64, 96, 128), steps_range=(20, 120))


def test_request_generator_distributions():
    reqs = S.serving_requests(300, 20.0, seed=1, sides=(16, 24, 32, 40), steps_range=(500, 20000))
    arr = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.2 / 20.0          # Poisson process of rate 20/s
    sides = np.array([r.scenario.s.nx for r in reqs])
    frac = [np.mean(sides == s) for s in (16, 24, 32, 40)]
    assert np.allclose(frac, [0.4, 0.3, 0.2, 0.1], atol=0.07)
    steps = np.array([r.steps for r in reqs])
    assert steps.min() >= 500 and steps.max() <= 20000
    lg = np.log(steps)
    assert abs(lg.mean() - 0.5 * (np.log(500) + np.log(20000))) < 0.25   # log-uniform


def test_summary_percentiles():
    recs = [serving.Record(i, 100, 10, arrival=0.0, finished=float(i + 1)) for i in range(100)]
    s = serving.summarize(recs, makespan=100.0)
    assert s["p50_latency_s"] == pytest.approx(50.5)
    assert s["p99_latency_s"] == pytest.approx(99.01)
    assert s["gcell_steps_per_s"] == pytest.approx(100 * 100 * 10 / 100.0 / 1e9)


def alone_state(req):
    sc = req.scenario
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(req.steps)
    st = e.get_state_f64()
    e.close()
    return st


@pytest.mark.gpu
def test_served_results_equal_alone():
    reqs = small_stream()
    recs, summ, states = serving.serve(reqs, block=16, keep_states=True, graph_steps=16)
    assert summ["requests"] == len(reqs)
    for r in reqs:
        assert np.array_equal(states[r.rid], alone_state(r)), f"request {r.rid} differs in the mix"
    for rec in recs:
        assert rec.finished >= rec.admitted >= rec.arrival - 1e-9


@pytest.mark.gpu
def test_budget_limits_admission():
    reqs = small_stream(n=4, seed=4)
    budget = max(estimate_device_bytes(r.scenario.s, r.scenario.boundaries) for r in reqs) + 1
    recs, summ, states = serving.serve(reqs, hbm_budget=budget, block=16, keep_states=True)
    recs = sorted(recs, key=lambda r: r.admitted)
    for a, b in zip(recs, recs[1:]):
        assert b.admitted >= a.finished - 1e-9        # one resident scenario at a time
    for r in reqs:
        assert np.array_equal(states[r.rid], alone_state(r))


@pytest.mark.gpu
def test_one_at_a_time_is_fifo():
    reqs = small_stream(n=4, seed=5)
    recs, summ = serving.one_at_a_time(reqs)
    recs = sorted(recs, key=lambda r: r.arrival)
    for a, b in zip(recs, recs[1:]):
        assert b.admitted == pytest.approx(max(b.arrival, a.finished))
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0


@pytest.mark.gpu
def test_srpt_results_equal_alone_and_short_requests_go_first():
    """The shortest-remaining-work policy pauses scenarios between blocks; every result is
    still bit-identical to the request run alone, and with room for one scenario at a time
    (gpu_cells 1) the requests complete in order of their work."""
    reqs = [S.Request(r.rid, 0.0, r.scenario, r.steps) for r in small_stream(n=6, seed=6)]   # all at once
    recs, summ, states = serving.serve(reqs, block=16, keep_states=True, policy="srpt", gpu_cells=1)
    assert summ["requests"] == len(reqs)
    for r in reqs:
        assert np.array_equal(states[r.rid], alone_state(r)), f"request {r.rid} differs under srpt"
    work = {r.rid: r.scenario.s.nx * r.scenario.s.ny * r.steps for r in reqs}
    order = [rec.rid for rec in sorted(recs, key=lambda rec: rec.finished)]
    assert [work[i] for i in order] == sorted(work[i] for i in order)


@pytest.mark.gpu
def test_paused_scenario_resumes_unchanged():
    from celeris_amd import Batch
    reqs = small_stream(n=2, seed=7)
    b = Batch()
    ids = [b.add(r.scenario.s, r.scenario.boundaries, r.scenario.bottom, r.scenario.state) for r in reqs]
    b.step(10)
    b.set_paused(ids[0])
    b.step(25)                                   # only the second advances
    b.set_paused(ids[0], False)
    b.step(5)
    t0, t1 = b.time(ids[0]), b.time(ids[1])
    s0 = b.get_state_f64(ids[0])
    b.close()
    a = reqs[0].scenario
    e = ShallowWaterEngine(a.s, a.boundaries, a.bottom, a.state)
    e.timestep(15)
    assert np.array_equal(s0, e.get_state_f64()) and t0 == pytest.approx(e.time, rel=0, abs=0)
    assert t1 > t0
    e.close()


class _FakeEngine:
    """Stand-in for a resident scenario (CPU only): records its steps."""

    def __init__(self, s, boundaries, bottom, state):
        self.nx, self.ny = s.nx, s.ny
        self.bad = not np.isfinite(state).all()


class _FakeBatch:
    """Stand-in for cel_batch on the CPU: members advance by counting; a member whose initial
    state is not finite is reported by failed() after its first step."""

    def __init__(self, graph_steps=0):
        self.m, self.n, self.paused, self.log, self.stepped = {}, 1, set(), [], set()

    def add(self, s, boundaries, bottom, state):
        i = self.n
        self.n += 1
        self.m[i] = _FakeEngine(s, boundaries, bottom, state)
        return i

    def set_paused(self, i, p):
        (self.paused.add if p else self.paused.discard)(i)

    def step(self, k):
        run = sorted(i for i in self.m if i not in self.paused)
        self.log.append((k, run))
        self.stepped.update(run)

    def sync(self):
        pass

    def failed(self):
        return [i for i, e in self.m.items() if e.bad and i in self.stepped]

    def remove(self, i):
        del self.m[i]

    def get_state_f64(self, i):
        return None

    def close(self):
        pass


class _Clock:
    """A clock that advances only when the loop sleeps or a block runs (1 ms per step)."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


def _fake_serve(monkeypatch, reqs, **kw):
    clk = _Clock()
    batches = []

    def make(graph_steps=0):
        b = _FakeBatch(graph_steps)
        orig = b.step

        def step(k):
            orig(k)
            clk.t += 1e-3 * k
        b.step = step
        batches.append(b)
        return b
    monkeypatch.setattr(serving, "Batch", make)
    monkeypatch.setattr(serving, "estimate_device_bytes", lambda s, bnd: s.nx * s.ny)
    out = serving.serve(reqs, clock=clk, sleep=clk.sleep, **kw)
    return out, batches[0]


def test_serve_logic_budget_srpt_and_eviction_on_cpu(monkeypatch):
    """The serving loop's host logic without a GPU: every request completes once, the HBM
    budget bounds the resident cells, "srpt" completes requests in order of their work when one
    runs at a time, and a request whose state is not finite is evicted as failed."""
    reqs = S.serving_requests(8, 50.0, seed=9, sides=(16, 24, 32, 40), steps_range=(20, 200))
    (recs, summ, _), b = _fake_serve(monkeypatch, reqs, block=16, hbm_budget=2 * 40 * 40)
    assert summ["requests"] == 8 and summ["failed"] == 0
    assert sorted(r.rid for r in recs) == list(range(8))
    for rec in recs:
        assert rec.finished >= rec.admitted >= rec.arrival - 1e-12
    # at most the budget's worth of cells resident: never more than two 40^2 grids
    assert all(len(run) <= 8 for _, run in b.log)
    # srpt with room for one grid at a time: completion in order of work
    at_once = [S.Request(r.rid, 0.0, r.scenario, r.steps) for r in reqs]
    (recs, summ, _), b = _fake_serve(monkeypatch, at_once, block=16, policy="srpt", gpu_cells=1)
    work = {r.rid: r.scenario.s.nx * r.scenario.s.ny * r.steps for r in reqs}
    order = [rec.rid for rec in sorted(recs, key=lambda rec: rec.finished)]
    assert [work[i] for i in order] == sorted(work.values())
    assert all(len(run) == 1 for _, run in b.log)
    # a request with a NaN in its state: evicted after its first block, the others finish
    bad = S.small(nx=20, ny=20, seed=3)
    bad.state[0, 5, 5] = np.nan
    mixed = at_once[:3] + [S.Request(99, 0.0, bad, 500)]
    (recs, summ, _), b = _fake_serve(monkeypatch, mixed, block=16)
    assert summ["failed"] == 1 and summ["requests"] == 3
    failed = [r for r in recs if r.failed]
    assert failed[0].rid == 99 and failed[0].finished < max(r.finished for r in recs if not r.failed)


def test_blocks_run_to_the_next_completion_once_nothing_waits(monkeypatch):
    """While requests wait for admission, blocks are at most `block` steps long (one
    admission per block); once every request is resident, each block ends exactly at the next
    completion."""
    reqs = [S.Request(i, 0.0, S.small(nx=16 + 4 * i, ny=16, seed=i), 40 + 30 * i) for i in range(4)]
    (recs, summ, _), b = _fake_serve(monkeypatch, reqs, block=8)
    assert summ["requests"] == 4 and summ["failed"] == 0
    ks = [k for k, _ in b.log]
    admitted_all = next(j for j, (_, run) in enumerate(b.log) if len(run) == 4)
    assert all(k <= 8 for k in ks[:admitted_all])
    # after the last admission: one block per completion, each running to it
    rest = b.log[admitted_all:]
    assert len(rest) == 4
    assert [len(run) for _, run in rest] == [4, 3, 2, 1]
    # every request advanced exactly its steps (members were admitted in arrival order)
    for sid, r in enumerate(reqs, start=1):
        assert sum(k for k, run in b.log if sid in run) == r.steps
