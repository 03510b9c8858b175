"""The batched-ensemble and serving workloads at their own sizes (GPU).

Ensemble: the 64 members of the ensemble workload (sides log-uniform in [128, 1024], own
seabeds and forcing, compensated update) advanced together in 50-step captured blocks:
every member bit-identical to itself run alone with the plain engine.

Serving: requests of the serving distribution's sides (256, 512, 1024, 2048) with at least
500 steps, the whole backlog present at once: every result bit-identical to the request run
alone; a request whose state is not finite is evicted (recorded as failed) without
disturbing the others.
"""
import numpy as np
import pytest

from celeris_amd import Batch, ShallowWaterEngine, serving
from celeris_amd import scenarios as S

pytestmark = pytest.mark.gpu


def alone(sc, steps):
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(steps)
    st = e.get_state_f64()
    e.close()
    return st


def test_ensemble_64_members_graph_blocks_equal_alone():
    members = S.ensemble(64, seed=4, steps=100)
    sides = [(m.s.nx, m.s.ny) for m in members]
    assert min(min(s) for s in sides) >= 128 and max(max(s) for s in sides) <= 1024
    b = Batch(graph_steps=50)
    ids = [b.add(m.s, m.boundaries, m.bottom, m.state) for m in members]
    b.step(100)
    assert b.failed() == []
    got = {i: b.get_state_f64(i) for i in ids}
    b.close()
    for i, m in zip(ids, members):
        ref = alone(m, 100)
        assert np.isfinite(ref).all()
        assert np.array_equal(got[i], ref), f"{m.name} ({m.s.nx}x{m.s.ny}) differs in the batch"


def test_serving_requests_256_to_2048_equal_alone_and_failed_evicted():
    rng = np.random.default_rng(12)
    reqs = []
    for rid, (side, steps) in enumerate(((256, 900), (512, 700), (1024, 600), (2048, 500), (256, 650))):
        sc = S.ensemble_member(rng, nx=side, ny=side, steps=steps, name=f"req{rid}")
        reqs.append(S.Request(rid, 0.0, sc, steps))
    bad = S.ensemble_member(rng, nx=300, ny=200, steps=600, name="bad")
    bad.state[0, 100, 150] = np.nan
    reqs.append(S.Request(len(reqs), 0.0, bad, 600))
    recs, summ, states = serving.serve(reqs, block=50, keep_states=True)
    assert summ["requests"] == len(reqs) - 1 and summ["failed"] == 1
    failed = [r for r in recs if r.failed]
    assert [r.rid for r in failed] == [bad_rid for bad_rid in (len(reqs) - 1,)]
    assert failed[0].finished < max(r.finished for r in recs)    # evicted before the others completed
    for r in reqs[:-1]:
        assert np.array_equal(states[r.rid], alone(r.scenario, r.steps)), f"request {r.rid} differs in the mix"
