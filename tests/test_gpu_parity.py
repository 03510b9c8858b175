"""GPU parity: the HIP engine (through the C ABI) against the FP32 CPU oracle.

Bar: 1e-5 relative on eta, hu, hv (max |diff| / max |oracle| per field) with identical
wet/dry masks.  Scenarios without trigonometric forcing are also expected to agree
bit-for-bit, since both sides perform the same IEEE operations in the same order; sine and
irregular forcing call sinf on the device and libm on the host.
"""
import numpy as np
import pytest

from celeris_amd import (EAST, IRREGULARWAVES, NORTH, SINEWAVE, SOLID, SOUTH, SPONGE, UNIFORMTIMESERIES, WEST,
                         Boundary, ShallowWaterEngine)
from celeris_amd import scenarios as S
from oracle.binding import Oracle
from tests.parity import eta, rel_err, wet_mask

pytestmark = pytest.mark.gpu
TOL = 1e-5


def run_pair(sc, steps, dts=None):
    eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    orc = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    if dts is None:
        eng.timestep(steps)
        orc.step(steps)
    else:
        for dt in dts:
            eng.reset_timestep(dt)
            orc.set_dt(dt)
            eng.timestep(1)
            orc.step(1)
    a, b = eng.get_state(), orc.state()
    assert eng.time == pytest.approx(orc.time, rel=0, abs=0)
    assert eng.scheme == orc.scheme
    eng.close()
    return a, b


def check(sc, a, b, exact=False):
    swl = sc.s.still_water_elevation
    assert np.isfinite(b).all(), "oracle state not finite"
    assert np.isfinite(a).all(), "engine state not finite"
    err = rel_err(eta(a, swl), eta(b, swl))
    assert (wet_mask(a, sc.bottom) == wet_mask(b, sc.bottom)).all(), f"wet/dry masks differ; {err}"
    if exact:
        assert np.array_equal(a, b), f"not bit-identical: {err}"
    for k, v in err.items():
        assert v <= TOL, f"{k} relative error {v:.3g} > {TOL}: {err}"
    return err


def test_solid_soliton_bitexact():
    sc = S.small(nx=48, ny=40)
    a, b = run_pair(sc, 200)
    check(sc, a, b, exact=True)


def test_odd_sizes_sponge():
    bnd = [Boundary(SPONGE, 6), Boundary(SPONGE, 5), Boundary(SOLID, 2), Boundary(SPONGE, 7)]
    sc = S.small(nx=37, ny=29, seed=3, bnd=bnd)
    a, b = run_pair(sc, 150)
    check(sc, a, b, exact=True)


def test_walls_meeting_sponges():
    """Solid sides meeting sponges at every corner: mirrored ghost cells whose source is
    rewritten by a sponge in the same boundary launch."""
    bnd = [Boundary(SOLID, 2), Boundary(SPONGE, 6), Boundary(SPONGE, 5), Boundary(SOLID, 2)]
    sc = S.small(nx=50, ny=40, seed=11, bnd=bnd)
    a, b = run_pair(sc, 150)
    check(sc, a, b, exact=True)


def test_narrow_channel_two_boundary_passes():
    """A channel narrower than two strips: the boundaries run as two launches (mirrored
    cells, then the rest)."""
    bnd = [Boundary(SOLID, 2), Boundary(SPONGE, 8), Boundary(SOLID, 2), Boundary(SPONGE, 6)]
    sc = S.small(nx=9, ny=60, seed=12, bnd=bnd)
    a, b = run_pair(sc, 150)
    check(sc, a, b, exact=True)


def test_wide_grid_multiple_strips_and_bands():
    sc = S.small(nx=157, ny=70, seed=5)
    a, b = run_pair(sc, 120)
    check(sc, a, b, exact=True)


def test_sine_west_sponge_east():
    bnd = [Boundary(SINEWAVE, 2, amplitude=0.03, period=1.5, theta=0.2), Boundary(SPONGE, 10), Boundary(SOLID, 2),
           Boundary(SOLID, 2)]
    sc = S.small(nx=64, ny=32, seed=1, bnd=bnd, soliton=False)
    a, b = run_pair(sc, 300)
    check(sc, a, b)


def test_sine_all_sides():
    bnd = [Boundary(SINEWAVE, 2, amplitude=0.02, period=1.7, theta=0.1),
           Boundary(SINEWAVE, 3, amplitude=0.02, period=2.1, theta=3.0),
           Boundary(SINEWAVE, 2, amplitude=0.01, period=1.3, theta=1.4),
           Boundary(SINEWAVE, 2, amplitude=0.01, period=1.9, theta=-1.6)]
    sc = S.small(nx=40, ny=44, seed=2, bnd=bnd, soliton=False)
    a, b = run_pair(sc, 200)
    check(sc, a, b)


def test_irregular_west_and_south():
    rng = np.random.default_rng(7)
    ww = S.jonswap_waves(37, 0.05, 2.0, 3.3, rng)
    ws = S.jonswap_waves(5, 0.03, 1.5, 3.3, rng)
    bnd = [Boundary(IRREGULARWAVES, 2, waves=ww), Boundary(SPONGE, 8), Boundary(IRREGULARWAVES, 2, waves=ws),
           Boundary(SOLID, 2)]
    sc = S.small(nx=64, ny=48, seed=4, bnd=bnd, soliton=False)
    a, b = run_pair(sc, 200)
    check(sc, a, b)


def test_uniform_time_series():
    t = np.arange(0, 40, 0.05, dtype=np.float32)
    uts = np.stack([t, 0.02 * np.sin(2 * np.pi * t / 2.0), 0.05 * np.sin(2 * np.pi * t / 2.0), 0 * t], 1)
    bnd = [Boundary(UNIFORMTIMESERIES, 2, uts=uts), Boundary(SOLID, 2), Boundary(SOLID, 2), Boundary(SOLID, 2)]
    sc = S.small(nx=50, ny=30, seed=6, bnd=bnd, soliton=False)
    a, b = run_pair(sc, 250)
    check(sc, a, b, exact=True)


def test_corrector_steps():
    sc = S.small(nx=44, ny=36, seed=8, correction_steps=2)
    a, b = run_pair(sc, 120)
    check(sc, a, b, exact=True)


def test_adaptive_scheme_variable_dt():
    sc = S.small(nx=40, ny=40, seed=9, time_scheme=3)
    dt0 = sc.s.timestep
    dts = [np.float32(dt0 * (1.0 + 0.3 * np.sin(0.1 * k))) for k in range(120)]
    a, b = run_pair(sc, None, dts=dts)
    check(sc, a, b, exact=True)


def test_manning_friction_nonzero_swl():
    sc = S.small(nx=48, ny=32, seed=10, friction=0.025, is_manning=1, still_water=0.3)
    a, b = run_pair(sc, 150)
    check(sc, a, b)


def test_quadratic_friction_non_boussinesq():
    sc = S.small(nx=48, ny=32, seed=11, friction=0.01, is_manning=0, is_boussinesq=0)
    a, b = run_pair(sc, 150)
    check(sc, a, b, exact=True)


def test_plain_beach_2000_steps():
    """The plain 256x256 beach case for its full 2000 steps: sine forcing, walls and an east
    sponge."""
    sc = S.plain_beach()
    a, b = run_pair(sc, 2000)
    err = check(sc, a, b)
    print("plain_beach 2000 steps relative errors", err)


def test_long_lines_register_chunks():
    """x-lines of 2048 and 1500 cells (32 and 24 unknowns per lane) and 1100-cell y-lines."""
    for nx, ny in ((2048, 20), (1500, 24), (24, 1100)):
        sc = S.small(nx=nx, ny=ny, seed=13)
        a, b = run_pair(sc, 40)
        check(sc, a, b, exact=True)


def test_lines_longer_than_a_wave_chunk():
    """2600-cell x-lines: 16 unknowns per thread, four waves per line."""
    sc = S.small(nx=2600, ny=12, seed=14)
    a, b = run_pair(sc, 30)
    check(sc, a, b, exact=True)


def test_longest_lines_and_smallest_grids():
    """Lines at the library's maximum (CEL_MAX_LINE = 8192 unknowns: the (16, 8) block shape,
    all 512 threads' chunks full) in either direction, a 4097-cell line (one past the (16, 4)
    shape's capacity), and the smallest grids (4 x 4 and 4 x 5 interior cells: one strip, one
    band, every interior cell next to a wall)."""
    for nx, ny in ((8192, 6), (6, 8192), (4097, 5), (4, 4), (4, 5)):
        sc = S.small(nx=nx, ny=ny, seed=15)
        a, b = run_pair(sc, 20)
        check(sc, a, b, exact=True)


def test_lines_above_the_maximum_rejected():
    from celeris_amd import CeleriError
    for nx, ny in ((8193, 8), (8, 8193), (3, 8)):
        sc = S.small(nx=max(nx, 4), ny=max(ny, 4), seed=16)
        sc.s.nx, sc.s.ny = nx, ny
        with pytest.raises(CeleriError) as e:
            ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        assert e.value.status == 1


def test_level0_coefficients_match_oracle():
    """A, B, C of both line directions, ghost rows/columns and margins included."""
    for kw in (dict(nx=37, ny=29, seed=3), dict(nx=40, ny=30, seed=4, is_boussinesq=0), dict(nx=33, ny=41, seed=5,
                                                                                          still_water=0.2)):
        sc = S.small(**kw)
        eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        ax, ay = eng.coefficients()
        eng.close()
        o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
        bx, by = o.coefficients()
        assert np.array_equal(ax, bx) and np.array_equal(ay, by)
