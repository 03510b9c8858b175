"""Reference input formats (celeris_amd.inputs) on the reference's own example files
(tests/golden/examples: the conical island, flume, diagonal solitary waves, irregular-wave
and uniform-time-series cases), the bathymetry sampling pinned by closed forms, the
hot-start path, and the logger's output formats."""
import os

import numpy as np
import pytest

from celeris_amd import (IRREGULARWAVES, PREDICTOR, SINEWAVE, SOLID, SPONGE, UNIFORMTIMESERIES, ShallowWaterEngine,
                         initial_state)
from celeris_amd import inputs as I

EX = os.path.join(os.path.dirname(__file__), "golden", "examples")


def test_parse_cml_unquoted_attributes():
    root = I.parse_cml(open(os.path.join(EX, "conical_island.cml")).read())
    assert root.name == "Experiment"
    p = root.child("model").child("parameters")
    assert p.attrs == {"epsilon": "5e-8", "Theta": "2", "correctionStepsNum": "2", "timestep": "0.005"}
    assert root.child("gridSize").attrs == {"nx": "201", "ny": "201"}
    assert root.child("bathymetryFilePath").text.strip() == "conical_island_400x400.cbf"
    assert [c.name for c in root.child("logData").children] == ["logPath", "range", "gauges"]


def test_conical_island_experiment():
    ex = I.load_experiment(os.path.join(EX, "conical_island.cml"))
    s = ex.s
    assert (s.nx, s.ny, s.width, s.length) == (201, 201, 30.0, 30.0)
    assert s.correction_steps_num == 2 and s.theta == 2.0 and s.is_manning == 1 and s.is_boussinesq == 1
    assert s.timestep == pytest.approx(0.005) and s.time_scheme == PREDICTOR
    assert [b.type for b in ex.boundaries] == [SPONGE, SPONGE, SOLID, SOLID]
    assert [b.width for b in ex.boundaries] == [20, 20, 2, 2]
    assert ex.solitons == [(pytest.approx(0.058), 0.0, pytest.approx(1e30), 5.0, 15.0)]
    assert ex.log.gauges[:2] == [(229, 302), (249, 302)] and len(ex.log.gauges) == 7
    assert ex.log.ranges == [("island", (228, 228), (374, 374))]
    # the island: flat -0.32 m offshore, a cone rising above the water at the centre
    BA = ex.bottom[2]
    assert BA.min() == pytest.approx(-0.32, abs=1e-6)
    assert BA[103, 103] > 0.0 and BA[5, 5] == pytest.approx(-0.32, abs=1e-6)
    assert np.isfinite(ex.state).all() and ex.state[0].max() > 0.05


def test_cbf_and_sampling_closed_forms():
    nx, ny, z = I.read_cbf(os.path.join(EX, "flume.cbf"))
    assert z.shape == (ny, nx) and z.dtype == np.float32
    # bilinear sampling is exact on a bilinear field and clamps at the edges
    zz = (np.add.outer(0.25 * np.arange(9), 0.5 * np.arange(13))).astype(np.float32)
    c = I.sample_corners(zz, 20, 10, 12.0, 8.0)
    x = ((np.arange(25, dtype=np.float32) - 2.5) / 19 * 12.0)
    y = ((np.arange(15, dtype=np.float32) - 2.5) / 9 * 8.0)
    gx = np.clip(x / 12.0 * 12, 0, 12)
    gy = np.clip(y / 8.0 * 8, 0, 8)
    exact = 0.25 * gy[:, None] + 0.5 * gx[None, :]
    np.testing.assert_allclose(c, exact, atol=2e-6)
    assert (c[:, 0] == c[:, 1]).all() and (c[0] == c[1]).all()        # clamped ghost corners


def test_irregular_and_time_series_tables():
    t = I.read_irregular_table(os.path.join(EX, "irrWavesWest.txt"))
    assert t.shape == (2, 4)
    np.testing.assert_allclose(t, [[0.04, 2, 0.785, 0.0], [0.04, 2, -0.785, 0.0]], rtol=1e-7)
    u = I.read_time_series(os.path.join(EX, "UTS_west.txt"))
    assert u.shape[1] == 4 and u.shape[0] > 300
    np.testing.assert_allclose(u[1], [0.05, 0.00782172, 0.01385836, 0.0], rtol=1e-6)
    assert (np.diff(u[:, 0]) > 0).all()
    ex = I.load_experiment(os.path.join(EX, "irregular.cml"))
    assert IRREGULARWAVES in [b.type for b in ex.boundaries]
    ex = I.load_experiment(os.path.join(EX, "uniform_time_series.cml"))
    assert UNIFORMTIMESERIES in [b.type for b in ex.boundaries]


def test_flume_and_diagonal_waves():
    ex = I.load_experiment(os.path.join(EX, "flume.cml"))
    assert (ex.s.nx, ex.s.ny) == (1600, 5) and [b.width for b in ex.boundaries] == [40, 40, 2, 2]
    assert ex.log.ranges[0][1:] == ((0, 3), (1603, 3))
    ex = I.load_experiment(os.path.join(EX, "diagSolitary.cml"))
    assert len(ex.solitons) == 4 and ex.s.still_water_elevation == 2.0
    assert ex.bottom[2].max() == 0.0 and ex.bottom[2].min() == 0.0        # "N/A": flat bed at 0
    thetas = sorted(round(np.degrees(t[1])) for t in ex.solitons)
    assert thetas == [-135, -45, 45, 135]


def test_hot_start_file(tmp_path):
    ex = I.load_experiment(os.path.join(EX, "diagSolitary.cml"))
    nx, ny = 20, 16
    s = ex.s
    s.nx, s.ny = nx, ny
    bottom = np.zeros((3, ny + 4, nx + 4), np.float32)
    hot = np.zeros((3, ny + 4, nx + 4), np.float32)
    rng = np.random.default_rng(0)
    hot[0] = 2.0 + rng.uniform(-0.01, 0.01, hot[0].shape)
    hot[1] = rng.uniform(-0.1, 0.1, hot[1].shape)
    with open(tmp_path / "hot.txt", "w") as f:
        for j in range(ny + 4):
            for i in range(nx + 4):
                f.write(f"{i} {j} {float(hot[0, j, i])!r} {float(hot[1, j, i])!r} {float(hot[2, j, i])!r} 0\n")
    h = I.read_hot_start(str(tmp_path / "hot.txt"), nx, ny)
    np.testing.assert_array_equal(h, hot)
    st = I.hot_start_state(s, bottom, h, [])
    np.testing.assert_array_equal(st, hot)               # no solitons: the file's state as is
    sol = [(0.1, 0.0, 1e30, 5.0, 8.0)]
    base = initial_state(s, bottom, sol) - initial_state(s, bottom, [])
    np.testing.assert_allclose(I.hot_start_state(s, bottom, h, sol) - hot, base, atol=1e-6)


@pytest.mark.gpu
def test_conical_island_runs_like_the_oracle(tmp_path):
    """The reference's conical island example through the drop-in: 201 x 201, Theta 2, two
    corrector steps per predictor, Manning friction, sponges W/E: within 1e-5 relative of
    the FP32 oracle with the same wet/dry mask."""
    from oracle.binding import Oracle
    from tests.parity import eta, rel_err, wet_mask
    ex = I.load_experiment(os.path.join(EX, "conical_island.cml"))
    e = ShallowWaterEngine(ex.s, ex.boundaries, ex.bottom, ex.state)
    o = Oracle(ex.s, ex.boundaries, ex.bottom, ex.state, "f32")
    e.timestep(152)
    o.step(152)
    a, b = e.get_state(), o.state()
    assert np.isfinite(a).all()
    err = rel_err(eta(a), eta(b))
    assert max(err.values()) < 1e-5, err
    assert (wet_mask(a, ex.bottom) == wet_mask(b, ex.bottom)).all()
    # the example's gauges and range lie outside its 205 x 205 grid (the reference reads past
    # its texture there): the logger leaves them out with a warning
    from celeris_amd.logger import Logger
    with pytest.warns(UserWarning):
        lg = Logger(e, ex.log, str(tmp_path))
    assert len(lg.skipped) == 8 and not lg.gauges and not lg.ranges
    # two Euler steps, then predictor + 2 corrector cycles: 152 steps in, the next step is a
    # predictor, and 30 more steps end ten cycles, the last with the final step; logStep 1
    # records every cycle
    ex.log.do_log, ex.log.log_step = True, 1
    ex.log.gauges = [(60, 100), (100, 100), (140, 100)]
    ex.log.ranges = [("island", (100, 100), (102, 101))]
    lg = Logger(e, ex.log, str(tmp_path))
    assert e.scheme == PREDICTOR
    lg.run(30)
    rows = [r for r in open(tmp_path / "gauges.txt").read().split("\n") if r]
    assert len(rows) == 10 * 3
    st = e.get_state()
    for (i, j), r in zip(ex.log.gauges, [r.split("\t") for r in rows[-3:]]):
        assert (int(r[0]), int(r[1])) == (i, j)
        assert float(r[2]) == pytest.approx(float(st[0, j, i]), rel=1e-5)
    assert len(open(tmp_path / "island.txt").read().strip().split("\n")) == 10 * 6
    assert len(open(tmp_path / "time.txt").read().split()) == 10
    txt = open(tmp_path / lg.dump_inundation()).read()
    assert txt.startswith("[nx] 201\n[ny] 201\n")
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cml,steps", [("flume.cml", 300), ("diagSolitary.cml", 150), ("irregular.cml", 150),
                                       ("uniform_time_series.cml", 150)])
def test_reference_examples_run_like_the_oracle(cml, steps):
    """The reference's other example experiments through the drop-in (sponge flume with a
    soliton, four diagonal solitons in a walled basin, irregular waves and uniform time
    series on all four sides of the conical island): within 1e-5 of the FP32 oracle with
    the same wet/dry mask."""
    from oracle.binding import Oracle
    from tests.parity import eta, rel_err, wet_mask
    ex = I.load_experiment(os.path.join(EX, cml))
    e = ShallowWaterEngine(ex.s, ex.boundaries, ex.bottom, ex.state)
    o = Oracle(ex.s, ex.boundaries, ex.bottom, ex.state, "f32")
    e.timestep(steps)
    o.step(steps)
    assert e.time == o.time
    a, b = e.get_state(), o.state()
    e.close()
    assert np.isfinite(a).all()
    swl = ex.s.still_water_elevation
    err = rel_err(eta(a, swl), eta(b, swl))
    assert max(err.values()) < 1e-5, err
    assert (wet_mask(a, ex.bottom) == wet_mask(b, ex.bottom)).all()
    assert np.abs(b[1:]).max() > 1e-3                   # the forcing or the waves moved water
