"""Captured multi-step blocks (hipGraph) against step-by-step launches of the same engine.

Every kernel parameter except the step clock is a function of the engine's discrete state,
so replays must reproduce step-by-step results bit for bit: across the Euler bootstrap,
with corrector steps (device copies inside the graph), with every time-dependent boundary
(sine, irregular waves, uniform time series read from the clock), with block lengths that
do not divide the ping-pong/ring period, and across a dt change (which drops the graphs).
"""
import numpy as np
import pytest

from celeris_amd import (IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, UNIFORMTIMESERIES, Boundary,
                         ShallowWaterEngine)
from celeris_amd import scenarios as S

pytestmark = pytest.mark.gpu


def forced_case(seed=21, correction_steps=0, nx=72, ny=56):
    rng = np.random.default_rng(seed)
    waves = np.stack([rng.uniform(0.005, 0.02, 6), rng.uniform(1.0, 3.0, 6), rng.uniform(-0.4, 0.4, 6),
                      rng.uniform(0, 2 * np.pi, 6)], axis=1).astype(np.float32)
    t = np.arange(0, 40, 0.05, dtype=np.float32)
    uts = np.stack([t, 0.02 * np.sin(t), 0.01 * np.cos(t), 0 * t], axis=1).astype(np.float32)
    bnd = [Boundary(IRREGULARWAVES, 2, waves=waves), Boundary(SPONGE, 10),
           Boundary(UNIFORMTIMESERIES, 2, uts=uts), Boundary(SINEWAVE, 2, amplitude=0.02, period=1.7, theta=0.3)]
    return S.small(nx=nx, ny=ny, seed=seed, bnd=bnd, correction_steps=correction_steps)


def run(sc, steps_list, graph_steps, dt_change=None):
    eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    eng.set_graph_steps(graph_steps)
    for i, n in enumerate(steps_list):
        if dt_change is not None and i == dt_change[0]:
            eng.reset_timestep(dt_change[1])
        eng.timestep(n)
    out = eng.get_state(), eng.time, eng.scheme
    eng.close()
    return out


@pytest.mark.parametrize("graph_steps,steps", [(7, [100]), (50, [130]), (6, [5, 36, 1, 18])])
def test_graph_blocks_bit_identical(graph_steps, steps):
    sc = forced_case()
    a, ta, sa = run(sc, steps, 0)
    b, tb, sb = run(sc, steps, graph_steps)
    assert np.isfinite(a).all()
    assert ta == tb and sa == sb
    assert np.array_equal(a, b)


def test_graph_blocks_with_corrector_steps():
    sc = forced_case(seed=22, correction_steps=2)
    a = run(sc, [90], 0)[0]
    b = run(sc, [90], 12)[0]
    assert np.array_equal(a, b)


def test_graph_blocks_across_dt_change():
    sc = forced_case(seed=23)
    dt2 = np.float32(0.8 * sc.s.timestep)
    a = run(sc, [40, 40, 40], 0, dt_change=(1, dt2))[0]
    b = run(sc, [40, 40, 40], 10, dt_change=(1, dt2))[0]
    assert np.array_equal(a, b)


def test_graph_replays_after_eager_work():
    """Graph blocks interleaved with eager steps and host synchronisations, replayed many
    times: each segment matches the step-by-step engine."""
    sc = forced_case(seed=24, nx=64, ny=64)
    e1 = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e2 = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e2.set_graph_steps(8)
    for n in (16, 3, 24, 1, 40):
        e1.timestep(n)
        e2.timestep(n)
        assert np.array_equal(e1.get_state(), e2.get_state())
        e2.stats()
    e1.close()
    e2.close()


def test_graph_steps_bounds():
    sc = S.small(nx=24, ny=24)
    eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    with pytest.raises(RuntimeError):
        eng.set_graph_steps(-1)
    with pytest.raises(RuntimeError):
        eng.set_graph_steps(4096)
    eng.close()
