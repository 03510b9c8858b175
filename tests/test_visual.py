"""Visual outputs of the step (compute.hlsl:437-457, 535-540: inundation record, slope angle,
whitewater, normals), addSolitaryWave (engine.cpp:2154-2220) and the logger's schedule.

CPU tests pin the oracle with closed forms: a linear surface reconstructs exactly, so the
mean face depth is w - B, the slope deviation is |dw/dx| and the normal is
(-b, -c, 1)/sqrt(1 + b^2 + c^2); the whitewater decays by |rate|^dt per step; an added
solitary wave on a lake at rest equals the initial-condition soliton of the same
parameters, its volume per crest metre is 2H/K and its momentum is C eta.  GPU tests
compare the engine with the oracle.
"""
import numpy as np
import pytest

from celeris_amd import SOLID, Boundary, initial_state, settings
from celeris_amd import scenarios as S
from oracle.binding import Oracle

G = float(np.float32(9.81))


def linear_case(nx=24, ny=20, dx=0.5, b=0.02, c=-0.01, depth=1.0, theta=1.0):
    """Flat bed at -depth, w = b x + c y over the whole ghost-inclusive grid (x = i dx)."""
    W, L = dx * (nx - 1), dx * (ny - 1)
    s = settings(nx, ny, W, L, 0.01, theta=theta)
    X, Y = nx + 4, ny + 4
    bottom = np.full((3, Y, X), -depth, np.float32)
    xs = np.arange(X) * np.float32(W / (nx - 1))
    ys = np.arange(Y) * np.float32(L / (ny - 1))
    state = np.zeros((3, Y, X), np.float32)
    state[0] = (b * xs[None, :] + c * ys[:, None]).astype(np.float32)
    return s, [Boundary(SOLID, 2) for _ in range(4)], bottom, state


def region_mask(nx, ny, lo=1, hi_off=2):
    m = np.zeros((ny + 4, nx + 4), bool)
    m[lo:ny + 4 - hi_off, lo:nx + 4 - hi_off] = True
    return m


def test_aux_first_step_closed_forms():
    """After one step the auxiliary texture holds Pass1 of the initial state: depth = w - B,
    slope = atan(|b|) (the x-slope dominates), whitewater = |rate|^dt where the slope
    angle exceeds the threshold, zero outside Pass2's cells."""
    nx, ny, b, c = 24, 20, 0.02, -0.01
    s, bnd, bottom, state = linear_case(nx, ny, b=b, c=c)
    o = Oracle(s, bnd, bottom, state, "f64")
    o.set_visual(0.01, 0.1)
    o.step(1)
    aux = o.aux()
    m = region_mask(nx, ny)
    np.testing.assert_allclose(aux[0][m], (state[0] - bottom[2])[m], rtol=0, atol=1e-6)
    np.testing.assert_allclose(aux[1][m], np.arctan(abs(b)), rtol=1e-4)
    rate = float(np.float32(1.0 / (1 + 2 * np.float32(0.1))))
    np.testing.assert_allclose(aux[2][m], rate ** float(np.float32(0.01)), rtol=1e-12)
    assert (aux[:, ~m] == 0).all()
    # above the slope angle nothing breaks
    o2 = Oracle(s, bnd, bottom, state, "f64")
    o2.set_visual(0.5, 0.1)
    o2.step(1)
    assert (o2.aux()[2] == 0).all()


def test_aux_whitewater_decay_and_record_monotone():
    nx, ny = 24, 20
    s, bnd, bottom, state = linear_case(nx, ny, b=0.02, c=0.0)
    o = Oracle(s, bnd, bottom, state, "f64")
    o.set_visual(0.0, 0.3)
    o.step(1)
    first = o.aux()
    # nothing breaks any more: pure decay.  The Pass1 at the end of step 1 already ran with
    # the old threshold (whitewater 1 again), so after 6 steps the factor is applied 5 times.
    o.set_visual(10.0, 0.3)
    prev = first
    for k in range(1, 6):
        o.step(1)
        a = o.aux()
        assert (a[0] >= prev[0]).all()                       # a running maximum
        prev = a
    m = region_mask(nx, ny)
    f = float(np.float32(1.0 / (1 + 2 * np.float32(0.3)))) ** float(np.float32(0.01))
    np.testing.assert_allclose(prev[2][m], f ** 5, rtol=1e-12)


def test_normals_closed_form():
    nx, ny, b, c = 24, 20, 0.05, -0.03
    s, bnd, bottom, state = linear_case(nx, ny, b=b, c=c)
    o = Oracle(s, bnd, bottom, state, "f64")
    o.refresh()
    n = o.normals()
    m = region_mask(nx, ny, 1, 1)
    ln = np.sqrt(1 + b * b + c * c)
    for k, v in enumerate((-b / ln, -c / ln, 1 / ln)):
        np.testing.assert_allclose(n[k][m], v, rtol=1e-5, atol=1e-7)


def lake(nx=64, ny=8, dx=0.5, depth=1.0, swl=0.0):
    W, L = dx * (nx - 1), dx * (ny - 1)
    s = settings(nx, ny, W, L, 0.005, still_water_elevation=swl)
    bottom = np.full((3, ny + 4, nx + 4), swl - depth, np.float32)
    return s, [Boundary(SOLID, 2) for _ in range(4)], bottom, initial_state(s, bottom, [])


@pytest.mark.parametrize("theta_deg", [0.0, 30.0])
def test_added_wave_equals_initial_condition_soliton(theta_deg):
    """On a lake at rest the added wave is the initial-condition soliton (createSimTextures'
    solitaryWave) with the same parameters, bit for bit (FP32 oracle vs the C ABI)."""
    s, bnd, bottom, state = lake()
    H, x0, y0 = 0.1, 12.0, 1.5
    PI_F = np.float32(4.0) * np.arctan(np.float32(1.0))
    theta = float(np.float32(np.float32(theta_deg) * PI_F / np.float32(180.0)))
    o = Oracle(s, bnd, bottom, state, "f32")
    o.add_solitary_wave(H, theta_deg, x0, y0)
    ic = initial_state(s, bottom, [(H, theta, 1e30, x0, y0)])
    np.testing.assert_array_equal(o.state(), ic)


def test_added_wave_volume_and_celerity():
    """Volume per crest metre 2H/K with K = sqrt(3H / 4 d^3); hu = sqrt(g (H + d)) eta; the
    crest height H at x0 (FP64 oracle)."""
    dx, d, H = 0.05, 1.0, 0.05
    s, bnd, bottom, state = lake(nx=2400, ny=4, dx=dx, depth=d)
    x0 = 1200 * dx
    o = Oracle(s, bnd, bottom, state, "f64")
    o.add_solitary_wave(H, 0.0, x0, 0.0)
    st = o.state()
    eta = st[0][3]
    K = np.sqrt(0.75 * H / d ** 3)
    np.testing.assert_allclose(eta.sum() * dx, 2 * H / K, rtol=1e-6)
    assert eta[1200] == pytest.approx(H, rel=1e-7)
    C = np.sqrt(G * (H + d))
    big = eta > 1e-6
    np.testing.assert_allclose(st[1][3][big] / eta[big], C, rtol=1e-5)
    assert np.abs(st[2]).max() == 0.0


def test_bathymetry_dump_round_trip(tmp_path):
    """dumpBathymetryToFile's .cbf (engine.cpp:750-779) reads back as the interior BA."""
    from types import SimpleNamespace
    from celeris_amd.inputs import LogConfig, read_cbf
    from celeris_amd.logger import Logger
    sc = S.small(20, 14, soliton=False)
    lg = Logger(SimpleNamespace(nx=20, ny=14), LogConfig(), str(tmp_path), inundation=False)
    nx, ny, z = read_cbf(str(tmp_path / lg.dump_bathymetry(sc.bottom)))
    assert (nx, ny) == (20, 14)
    np.testing.assert_allclose(z, sc.bottom[2][2:16, 2:22], rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_engine_visual_outputs_match_oracle():
    from celeris_amd import ShallowWaterEngine
    from tests.parity import eta, rel_err
    sc = S.small(48, 40, correction_steps=1)
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    e.set_visual(True, 0.01, 0.2)
    o.set_visual(0.01, 0.2)
    e.timestep(60)
    o.step(60)
    a, b = e.aux(), o.aux()
    assert max(rel_err(eta(e.get_state()), eta(o.state())).values()) < 1e-5
    np.testing.assert_allclose(a[0], b[0], rtol=0, atol=1e-6)           # inundation record
    np.testing.assert_allclose(a[1], b[1], rtol=0, atol=2e-6)           # slope angle (device atanf)
    np.testing.assert_allclose(a[2], b[2], rtol=0, atol=2e-6)           # whitewater
    assert (b[2] > 0).any() and (b[0] > 0).any()
    o.refresh()
    np.testing.assert_allclose(e.normals(), o.normals(), rtol=0, atol=2e-6)
    # a hot start clears the record; disabling refuses the read-back
    e.set_state(sc.state)
    assert (e.aux() == 0).all()
    e.set_visual(False)
    with pytest.raises(Exception):
        e.aux()
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_engine_add_solitary_wave_matches_oracle(precision):
    """A wave added mid-run: the engine's next step sees the new state, as the oracle with
    a refreshed Pass1; both then run on within the parity bar."""
    from celeris_amd import ShallowWaterEngine
    from tests.parity import eta, rel_err, wet_mask
    sc = S.small(48, 40, soliton=False)
    sc.s.state_precision = precision
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    e.timestep(20)
    o.step(20)
    e.add_solitary_wave(0.08, 20.0, 6.0, 10.0)
    o.add_solitary_wave(0.08, 20.0, 6.0, 10.0, refresh=True)
    err = rel_err(eta(e.get_state()), eta(o.state()))
    assert max(err.values()) < 1e-6, err
    e.timestep(80)
    o.step(80)
    a, b = e.get_state(), o.state()
    err = rel_err(eta(a), eta(b))
    assert max(err.values()) < 1e-5, err
    assert (wet_mask(a, sc.bottom) == wet_mask(b, sc.bottom)).all()
    e.close()


@pytest.mark.gpu
def test_logger_schedule_and_dumps(tmp_path):
    """correctionStepsNum 1, logStep 2: records after steps 2, 6, 10 of 12 with the virtual
    time before each cycle's dt (dt, 3 dt, 5 dt), as engine.cpp:1488-1518 schedules them;
    gauges equal the state; the inundation dump is the auxiliary record."""
    from celeris_amd import ShallowWaterEngine
    from celeris_amd.inputs import LogConfig, read_cbf
    from celeris_amd.logger import Logger
    sc = S.small(32, 24, correction_steps=1)
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    cfg = LogConfig(do_log=True, log_step=2, ranges=[("strip", (3, 4), (6, 5))],
                    gauges=[(5, 6), (20, 12), (500, 3)])
    with pytest.warns(UserWarning):
        lg = Logger(e, cfg, str(tmp_path))
    ref = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    states = {}
    for k in range(1, 13):
        ref.timestep(1)
        states[k] = ref.get_state()
    lg.run(12)
    dt = np.float32(sc.s.timestep)
    times = [float(v) for v in open(tmp_path / "time.txt").read().split()]
    expect = [float(dt), float(np.float32(np.float32(dt + dt) + dt)),
              float(np.float32(np.float32(np.float32(np.float32(dt + dt) + dt) + dt) + dt))]
    np.testing.assert_allclose(times, expect, rtol=1e-5)
    assert len(open(tmp_path / "time-elapsed.txt").read().split()) == 3
    rows = [r.split("\t") for r in open(tmp_path / "gauges.txt").read().strip().split("\n")]
    assert len(rows) == 3 * 2
    for rec, step in enumerate((2, 6, 10)):
        for g, (i, j) in enumerate([(5, 6), (20, 12)]):
            r = rows[2 * rec + g]
            assert (int(r[0]), int(r[1])) == (i, j) and r[5] == "0"
            assert float(r[2]) == pytest.approx(float(states[step][0, j, i]), rel=1e-5, abs=1e-6)
    assert len(open(tmp_path / "strip.txt").read().strip().split("\n")) == 3 * 8
    name = lg.dump_inundation()
    lines = open(tmp_path / name).read().split("\n")
    assert lines[0] == "[nx] 32" and lines[1] == "[ny] 24"
    body = [l for l in lines[7:] if l]
    assert len(body) == 24 and len(body[0].rstrip("\t").split("\t")) == 33
    aux = e.aux()
    np.testing.assert_allclose([float(v) for v in body[0].rstrip("\t").split("\t")], aux[0][2, 2:35], rtol=1e-5)
    bname = lg.dump_bathymetry(sc.bottom)
    nx, ny, z = read_cbf(str(tmp_path / bname))
    assert (nx, ny) == (32, 24)
    np.testing.assert_allclose(z, sc.bottom[2][2:26, 2:34], rtol=1e-5)
    e.close()
    ref.close()
