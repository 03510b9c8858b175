"""Work decomposition (band height of the explicit stencil, line-solve block shape) and
the autotuner.  Every choice computes each cell and every elimination the same way, so
states must be bit-identical whatever the tuning: fixed choices, autotuned engines and
autotuned batches against default engines."""
import numpy as np
import pytest

from celeris_amd import (IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, STATE_COMPENSATED, Batch, Boundary, CeleriError,
                         ShallowWaterEngine)
from celeris_amd import scenarios as S

pytestmark = pytest.mark.gpu


def case(nx, ny, seed, comp=False):
    rng = np.random.default_rng(seed)
    waves = np.stack([rng.uniform(0.005, 0.02, 4), rng.uniform(1.0, 3.0, 4), rng.uniform(-0.4, 0.4, 4),
                      rng.uniform(0, 2 * np.pi, 4)], axis=1).astype(np.float32)
    bnd = [Boundary(IRREGULARWAVES, 2, waves=waves), Boundary(SPONGE, 9), Boundary(SOLID, 2),
           Boundary(SINEWAVE, 2, amplitude=0.02, period=1.3, theta=0.2)]
    sc = S.small(nx=nx, ny=ny, seed=seed, bnd=bnd)
    if comp:
        sc.s.state_precision = STATE_COMPENSATED
    return sc


def run(sc, steps, tuning=None, autotune=False):
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    if tuning:
        e.set_tuning(*tuning)
    e.timestep(steps // 2)
    if autotune:
        t = e.autotune(repeats=3)
        assert t["band_rows"] in (2, 4, 8, 12, 16) and t["explicit_us"] > 0
    e.timestep(steps - steps // 2)
    st = e.get_state_f64()
    e.close()
    return st


@pytest.mark.parametrize("tuning", [(4, 1, 2, 2, 2), (16, 2, 1, 1, 2), (8, 1, 2, 4, 1), (16, 4, 4, 8, 2), (12, 2, 1, 2, 1),
                                    (2, 2, 1, 2, 1), (2, 2, 1, 2, 1, False)])
def test_fixed_tunings_bit_identical(tuning):
    sc = case(100, 70, 41)
    assert np.array_equal(run(sc, 60), run(sc, 60, tuning=tuning))


@pytest.mark.parametrize("nx,ny,comp", [(180, 90, False), (600, 40, True), (30, 300, False)])
def test_autotuned_engine_bit_identical(nx, ny, comp):
    sc = case(nx, ny, 42, comp)
    assert np.array_equal(run(sc, 50), run(sc, 50, autotune=True))


def test_invalid_tunings_rejected():
    sc = case(300, 40, 43)
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    for bad in [(5, 8, 1, 1, 1), (8, 1, 1, 1, 1), (8, 3, 2, 1, 1), (8, 16, 16, 1, 1)]:
        with pytest.raises(CeleriError):
            e.set_tuning(*bad)
    e.close()


def test_autotuned_batch_bit_identical():
    scs = [case(70, 50, 44), case(300, 120, 45, comp=True), case(40, 260, 46)]
    b = Batch()
    ids = [b.add(sc.s, sc.boundaries, sc.bottom, sc.state) for sc in scs]
    b.step(20)
    rows = b.autotune(repeats=3, tune_lines=True)
    assert rows in (4, 8, 16)
    b.step(30)
    for sid, sc in zip(ids, scs):
        assert np.array_equal(b.get_state_f64(sid), run(sc, 50))
    b.close()


def shoreline(seed):
    """A beach with wetting and drying and a soliton: divisions and roots see depths from
    zero up."""
    return S.small(nx=90, ny=64, seed=seed, soliton=True)


@pytest.mark.parametrize("make", [lambda: case(100, 70, 47), lambda: case(120, 60, 48, comp=True),
                                  lambda: shoreline(49)])
def test_fast_arith_bit_identical(make):
    """The explicit kernel's shortened exact division and square root give the bits of the
    library operations: the engine that uses them equals one that is denied them."""
    sc = make()
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    t = e.tuning()
    assert t["fast_arith"] == 1
    e.timestep(80)
    a = e.get_state_f64()
    e.close()
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.set_tuning(t["band_rows"], t["x_chunk"], t["x_waves"], t["y_chunk"], t["y_waves"], fast_arith=False)
    assert e.tuning()["fast_arith"] == 0
    e.timestep(80)
    b = e.get_state_f64()
    e.close()
    assert np.array_equal(a, b)


def test_fast_arith_ruled_out_by_mesh_and_settings():
    """A depth of 1e-30 below the sea level (K1's d^2 underflows), an epsilon below 2^-96 or
    a grid spacing below 2^-10 keep the plain kernel; a new bottom re-decides."""
    sc = shoreline(50)
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    assert e.tuning()["fast_arith"] == 1
    nb = sc.bottom.copy()
    nb[2, 30, 40] = np.float32(sc.s.still_water_elevation) - np.float32(1e-30)
    e.set_bottom(nb, keep_depth=True)
    assert e.tuning()["fast_arith"] == 0
    e.timestep(5)
    e.set_bottom(sc.bottom, keep_depth=True)
    assert e.tuning()["fast_arith"] == 1
    e.close()
    sc = shoreline(51)
    sc.s.epsilon = 1e-30
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    assert e.tuning()["fast_arith"] == 0
    e.close()
    nx, ny = 40, 30
    s = S.settings(nx, ny, 5e-4 * (nx - 1), 5e-4 * (ny - 1), 1e-5)
    bottom = np.full((3, ny + 4, nx + 4), -0.01, np.float32)
    e = ShallowWaterEngine(s, [Boundary(SOLID, 2)] * 4, bottom, S.initial_state(s, bottom, []))
    assert e.tuning()["fast_arith"] == 0
    e.close()


@pytest.mark.parametrize("nx,ny,tuning", [(256, 128, (8, 4, 1, 2, 1)), (128, 512, (12, 1, 2, 8, 1)),
                                          (512, 64, (8, 8, 1, 1, 1)), (256, 129, (2, 4, 1, 4, 1))])
def test_lines_filling_their_blocks_match_oracle(nx, ny, tuning):
    # lines of exactly 64 W C unknowns take the solver's full-line form (no tests against the
    # line's end); the same eliminations, so the FP32 oracle's bits
    from oracle.binding import Oracle
    sc = S.small(nx=nx, ny=ny, seed=47, soliton=True,
                 bnd=[Boundary(SPONGE, 9), Boundary(SOLID, 2), Boundary(SOLID, 2), Boundary(SPONGE, 7)])
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.set_tuning(*tuning)
    e.timestep(30)
    a = e.get_state()
    e.close()
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    o.step(30)
    assert np.array_equal(a, o.state())
