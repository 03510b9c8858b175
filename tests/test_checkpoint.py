"""Checkpoint and resume (cel_engine_checkpoint / cel_engine_restore): an engine restored from
a checkpoint continues bit-identically to the engine it was taken from -- state, time and
scheme -- through the Euler bootstrap, AB3 with every boundary type (time-series cursor
included), corrector steps (the ring's corrector semantics), variable time steps (the dt
history), the compensated and the FP64 states and the visual record.  The reference's hot
start keeps only (w, hu, hv) and restarts with Euler steps; that path is cel_engine_set_state."""
import numpy as np
import pytest

from celeris_amd import (IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, STATE_COMPENSATED, STATE_FP64, UNIFORMTIMESERIES,
                         Boundary, CeleriError, ShallowWaterEngine)
from celeris_amd import scenarios as S

pytestmark = pytest.mark.gpu


def every_boundary(seed, **kw):
    rng = np.random.default_rng(seed)
    waves = np.stack([rng.uniform(0.005, 0.02, 6), rng.uniform(1.0, 3.0, 6), rng.uniform(-0.4, 0.4, 6),
                      rng.uniform(0, 2 * np.pi, 6)], axis=1).astype(np.float32)
    t = np.arange(0, 30, 0.07, dtype=np.float32)
    uts = np.stack([t, 0.02 * np.sin(t), 0.01 * np.cos(t), 0 * t], axis=1).astype(np.float32)
    bnd = [Boundary(IRREGULARWAVES, 2, waves=waves), Boundary(SPONGE, 9), Boundary(UNIFORMTIMESERIES, 2, uts=uts),
           Boundary(SINEWAVE, 2, amplitude=0.02, period=1.3, theta=0.2)]
    return S.small(nx=90, ny=70, seed=seed, bnd=bnd, **kw)


def resume_matches(sc, before, after, dts=None, visual=False):
    a = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    if visual:
        a.set_visual(True)

    def advance(e, n, offset):
        if dts is None:
            e.timestep(n)
        else:
            for k in range(n):
                e.reset_timestep(dts[(offset + k) % len(dts)])
                e.timestep(1)

    advance(a, before, 0)
    ck = a.checkpoint()
    advance(a, after, before)
    b = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    if visual:
        b.set_visual(True)
    b.restore(ck)
    advance(b, after, before)
    assert np.array_equal(a.get_state_f64(), b.get_state_f64())
    assert a.time == b.time and a.scheme == b.scheme
    if visual:
        assert np.array_equal(a.aux(), b.aux())
    a.close()
    b.close()


@pytest.mark.parametrize("before", [1, 2, 30])
def test_resume_every_boundary_type(before):
    resume_matches(every_boundary(51), before, 40)


def test_resume_corrector_steps():
    sc = S.small(nx=64, ny=56, seed=52)
    sc.s.correction_steps_num = 2
    for before in (3, 4, 5, 20):        # predictor and corrector positions of the cycle
        resume_matches(sc, before, 25)


def test_resume_variable_steps_and_visual_record():
    sc = every_boundary(53, time_scheme=3)
    resume_matches(sc, 17, 30, dts=[sc.s.timestep * f for f in (1.0, 0.8, 1.1, 0.9)], visual=True)


def test_resume_compensated_and_fp64():
    for prec in (STATE_COMPENSATED, STATE_FP64):
        sc = every_boundary(54)
        sc.s.state_precision = prec
        resume_matches(sc, 12, 20)


def test_restore_rejects_another_engine():
    sc = S.small(nx=40, ny=36, seed=55)
    a = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    a.timestep(3)
    ck = a.checkpoint()
    a.close()
    other = S.small(nx=44, ny=36, seed=55)
    b = ShallowWaterEngine(other.s, other.boundaries, other.bottom, other.state)
    with pytest.raises(CeleriError) as e:
        b.restore(ck)
    assert e.value.status == 1
    with pytest.raises(CeleriError):
        b.restore(ck[:100])
    b.close()


def test_restore_rejects_a_damaged_step_state():
    """A checkpoint whose step scalars are out of range (a buffer or ring slot, the scheme,
    a time-series row before the first) is refused before anything changes: the engine
    continues as if the restore had not been tried."""
    sc = every_boundary(56)
    a = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    a.timestep(5)
    ck = a.checkpoint()
    header = 12 * 4                      # magic + eleven ints
    scal = ck[header:header + 19 * 8].view(np.float64)
    for k, value in [(0, 2.0), (1, 3.0), (2, 0.5), (1, float(scal[2])), (4, 7.0), (8, float("nan")), (11, 0.0)]:
        bad = ck.copy()
        bad[header:header + 19 * 8].view(np.float64)[k] = value
        with pytest.raises(CeleriError) as e:
            a.restore(bad)
        assert e.value.status == 1
    a.timestep(7)
    b = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    b.timestep(12)
    assert np.array_equal(a.get_state_f64(), b.get_state_f64()) and a.time == b.time
    a.close()
    b.close()
