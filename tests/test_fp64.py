"""FP64 variant (CEL_STATE_FP64): every kernel and host-side table in double precision,
against the FP64 oracle.  Both sides perform the same IEEE double operations in the same
order (correctly rounded division and square root), so scenarios without trigonometric
forcing agree to round-off (checked at 1e-12 relative); forcing uses the device's and the
host's double sin (1e-10).  The precision study compares FP32, compensated and FP64 runs
of a solitary wave on a beach with the FP64 oracle."""
import numpy as np
import pytest

from celeris_amd import (IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, STATE_COMPENSATED, STATE_FP32, STATE_FP64,
                         UNIFORMTIMESERIES, Batch, Boundary, CeleriError, ShallowWaterEngine)
from celeris_amd import scenarios as S
from oracle.binding import Oracle
from tests.parity import eta, rel_err, wet_mask

pytestmark = pytest.mark.gpu


def fp64(sc):
    sc.s.state_precision = STATE_FP64
    return sc


def pair(sc, steps, dts=None):
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    if dts is None:
        e.timestep(steps)
        o.step(steps)
    else:
        for dt in dts:
            e.reset_timestep(dt)
            o.set_dt(dt)
            e.timestep(1)
            o.step(1)
    a, b = e.get_state_f64(), o.state()
    assert e.time == pytest.approx(o.time, rel=1e-15)
    assert e.scheme == o.scheme
    e.close()
    return a, b


def check(sc, a, b, tol):
    swl = sc.s.still_water_elevation
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert (wet_mask(a, sc.bottom) == wet_mask(b, sc.bottom)).all()
    err = rel_err(eta(a, swl), eta(b, swl))
    assert max(err.values()) <= tol, err
    return err


@pytest.mark.parametrize("kw,steps", [(dict(nx=48, ny=40), 200), (dict(nx=44, ny=36, seed=8, correction_steps=2), 120),
                                      (dict(nx=61, ny=37, seed=3, bnd=[Boundary(SPONGE, 6), Boundary(SPONGE, 5),
                                                                        Boundary(SOLID, 2), Boundary(SPONGE, 7)]), 150),
                                      (dict(nx=130, ny=20, seed=5, friction=0.02, is_manning=1, still_water=0.2), 120)])
def test_fp64_matches_fp64_oracle(kw, steps):
    sc = fp64(S.small(**kw))
    a, b = pair(sc, steps)
    check(sc, a, b, 1e-12)


def test_fp64_variable_dt():
    sc = fp64(S.small(nx=40, ny=40, seed=9, time_scheme=3))
    dts = [np.float32(sc.s.timestep * (1.0 + 0.3 * np.sin(0.1 * k))) for k in range(100)]
    a, b = pair(sc, None, dts=dts)
    check(sc, a, b, 1e-12)


def test_fp64_forcing():
    rng = np.random.default_rng(7)
    waves = np.stack([rng.uniform(0.005, 0.02, 5), rng.uniform(1.0, 3.0, 5), rng.uniform(-0.4, 0.4, 5),
                      rng.uniform(0, 2 * np.pi, 5)], axis=1).astype(np.float32)
    t = np.arange(0, 40, 0.05, dtype=np.float32)
    uts = np.stack([t, 0.02 * np.sin(t), 0.01 * np.cos(t), 0 * t], axis=1).astype(np.float32)
    bnd = [Boundary(IRREGULARWAVES, 2, waves=waves), Boundary(SPONGE, 8), Boundary(UNIFORMTIMESERIES, 2, uts=uts),
           Boundary(SINEWAVE, 2, amplitude=0.02, period=1.5, theta=0.3)]
    sc = fp64(S.small(nx=72, ny=56, seed=21, bnd=bnd))
    a, b = pair(sc, 150)
    check(sc, a, b, 1e-10)


def test_fp64_graph_and_stats():
    sc = fp64(S.small(nx=64, ny=48, seed=12))
    e1 = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e2 = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e2.set_graph_steps(9)
    e1.timestep(60)
    e2.timestep(60)
    assert np.array_equal(e1.get_state_f64(), e2.get_state_f64())
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    o.step(60)
    g, r = e1.stats(), o.stats()
    assert g["sum_h"] == pytest.approx(r[0], rel=1e-12)
    assert g["max_cfl_rate"] == pytest.approx(r[7], rel=1e-12)
    e1.close()
    e2.close()


def test_fp64_rejected_in_batches():
    sc = fp64(S.small(nx=24, ny=24))
    b = Batch()
    with pytest.raises(CeleriError):
        b.add(sc.s, sc.boundaries, sc.bottom, sc.state)
    b.close()


def test_precision_study_solitary_beach():
    """A solitary wave (H/h = 0.3) running up a 1:50 beach: the FP64 engine tracks the FP64
    oracle to round-off; FP32 and compensated stay within 1e-3 of the wave height."""
    sc = S.solitary_channel(nx=384, ny=16, dx=0.25, steps=1500)
    H = 0.3
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    o.step(sc.steps)
    ref = o.state()
    errs = {}
    for name, prec in (("fp32", STATE_FP32), ("compensated", STATE_COMPENSATED), ("fp64", STATE_FP64)):
        sc.s.state_precision = prec
        e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        e.timestep(sc.steps)
        a = e.get_state_f64()
        e.close()
        d = (a[0] - ref[0])[2:-2, 2:-2]
        errs[name] = (np.abs(d).max() / H, np.sqrt((d ** 2).mean()) / H)
    print("eta error / H (Linf, L2):", {k: ("%.2e" % v[0], "%.2e" % v[1]) for k, v in errs.items()})
    assert errs["fp64"][0] < 1e-10
    assert errs["fp32"][0] < 1e-3 and errs["compensated"][0] < 1e-3


def _runup(state, bottom, row):
    """Highest bed elevation reached by water (h > 1e-3 m) along one row."""
    h = state[0][row] - bottom[2][row]
    wet = h > 1e-3
    return float(bottom[2][row][wet].max()) if wet.any() else float("nan")


@pytest.mark.slow
def test_precision_study_reduced_c3():
    """The precision case on a 2048 x 8 channel (dx 0.25 m, 1:50 beach to +0.5 m, solitary wave
    H/h = 0.3 from the west, walls all round), 24000 steps: the wave crosses the channel and
    runs up the beach.  Reports eta error (L-inf, L2, over H) against the FP64 oracle, the
    runup height and the volume drift of each variant."""
    sc = S.solitary_channel(nx=2048, ny=8, dx=0.25, steps=24000)
    H, row, chunk = 0.3, 6, 2000
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    v0 = o.stats()[0]
    ref_runup = 0.0
    for _ in range(sc.steps // chunk):
        o.step(chunk)
        ref_runup = max(ref_runup, _runup(o.state(), sc.bottom, row))
    ref = o.state()
    out = {"oracle": {"runup_m": ref_runup, "volume_drift": (o.stats()[0] - v0) / v0}}
    for name, prec in (("fp32", STATE_FP32), ("compensated", STATE_COMPENSATED), ("fp64", STATE_FP64)):
        sc.s.state_precision = prec
        e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        g0 = e.stats()["sum_h"]
        runup = 0.0
        for _ in range(sc.steps // chunk):
            e.timestep(chunk)
            runup = max(runup, _runup(e.get_state_f64(), sc.bottom, row))
        a = e.get_state_f64()
        d = (a[0] - ref[0])[2:-2, 2:-2]
        out[name] = {"eta_linf_over_H": float(np.abs(d).max() / H), "eta_l2_over_H": float(np.sqrt((d ** 2).mean()) / H),
                     "runup_m": runup, "volume_drift": (e.stats()["sum_h"] - g0) / g0}
        e.close()
    print("precision study (2048x8, 24000 steps):", out)
    assert out["fp64"]["eta_linf_over_H"] < 1e-9
    assert out["fp32"]["eta_linf_over_H"] < 1e-2 and out["compensated"]["eta_linf_over_H"] < 1e-2
    assert abs(out["fp64"]["runup_m"] - ref_runup) < 1e-9
