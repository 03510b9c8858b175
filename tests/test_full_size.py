"""Parity and invariants at the benchmark sizes: the 2048 x 2048 headline workload and the
4096 x 1024 precision channel.  The CPU oracle is fast enough (OpenMP) for a few dozen
steps at these sizes, so the engine is compared with it directly; longer runs check
size-independent properties (mass of a closed basin, a lake at rest)."""
import numpy as np
import pytest

from celeris_amd import SOLID, STATE_FP64, Boundary, ShallowWaterEngine, bottom_from_corners, initial_state, settings
from celeris_amd import scenarios as S
from oracle.binding import Oracle
from tests.parity import eta, rel_err, wet_mask

pytestmark = pytest.mark.gpu


def test_headline_2048_1000_steps_matches_oracle():
    """2048^2, 100 irregular-wave components on the west side, 20-cell east sponge, after 1000
    steps (Euler bootstrap, then AB3): within 1e-5 of the FP32 oracle (device sinf against the
    C library's) with the same wet/dry mask."""
    sc = S.headline()
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    o.step(1000)
    b = o.state()
    del o
    assert np.abs(b[1]).max() > 0                    # the waves have entered
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(1000)
    a = e.get_state()
    e.close()
    assert np.isfinite(a).all()
    err = rel_err(eta(a), eta(b))
    print("headline 2048^2, 1000 steps, relative errors", err)
    assert max(err.values()) < 1e-5, err
    assert (wet_mask(a, sc.bottom) == wet_mask(b, sc.bottom)).all()


def test_precision_channel_4096x1024_matches_oracle():
    """The precision case at its full size: FP32 engine == FP32 oracle bit for bit (no
    trigonometric forcing) after 200 steps; the FP64 engine equals the FP64 oracle to
    round-off after 1000 steps, with the solitary wave moved about 20 cells."""
    sc = S.solitary_channel()
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    e.timestep(200)
    o.step(200)
    assert np.array_equal(e.get_state(), o.state())
    e.close()
    del o
    sc.s.state_precision = STATE_FP64
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    e.timestep(1000)
    o.step(1000)
    b = o.state()
    np.testing.assert_allclose(e.get_state_f64(), b, rtol=0, atol=1e-11)
    e.close()
    # the crest has moved: the surface differs from the initial state by a good part of H
    assert np.abs(b[0] - sc.state[0].astype(np.float64)).max() > 0.05


def bowl(nx=2048, dx=1.0, seed=1):
    """A closed basin whose shore stays inside the domain: paraboloid bowl from -3 m to a dry
    rim, a Gaussian random field (std 0.3 m), an oblique solitary wave, walls all round."""
    W = L = dx * (nx - 1)
    X, Y = S.corner_coords(nx, nx, W, L)
    r2 = ((X - W / 2) ** 2 + (Y - L / 2) ** 2) / (0.45 * W) ** 2
    rng = np.random.default_rng(seed)
    z = -3.0 + 4.0 * r2 + S.gaussian_random_field(X.shape, 20.0, 0.3, rng)
    s = settings(nx, nx, W, L, S.cfl_dt(dx, 4.0), theta=1.3)
    bottom = bottom_from_corners(nx, nx, z.astype(np.float32))
    state = initial_state(s, bottom, [(0.3, 0.5, 1e6, 0.4 * W, 0.5 * L)])
    return s, [Boundary(SOLID, 2) for _ in range(4)], bottom, state


# The FP32 oracle's total depth (FP64 sum over the interior) of the bowl before and after 400
# steps, computed on the CPU (about 100 s with 8 threads): the volume grows by 8.8e-7 of
# itself through wetting and drying on the shore, in the FP64 oracle alike (8.8e-7).
BOWL_SUM_H_0 = 3030033.091669999
BOWL_SUM_H_400 = 3030035.751949297


def test_closed_bowl_2048_volume_budget():
    """The water never reaches the walls; its volume changes only through the scheme's
    wetting and drying.  After 400 steps at 2048^2 the engine's total depth equals the FP32
    oracle's to the order of the FP64 summation."""
    s, bnd, bottom, state = bowl()
    assert not ((state[0] - bottom[2])[2:-2, 2:-2][[0, -1], :] > 1e-4).any()
    e = ShallowWaterEngine(s, bnd, bottom, state)
    m0 = e.stats()["sum_h"]
    e.timestep(400)
    st = e.stats()
    a = e.get_state()
    e.close()
    assert np.isfinite(a).all() and np.abs(a[1]).max() > 0.1
    assert m0 == pytest.approx(BOWL_SUM_H_0, rel=1e-12)
    assert st["sum_h"] == pytest.approx(BOWL_SUM_H_400, rel=1e-11)
    assert abs(st["sum_h"] - m0) / m0 < 2e-6


def test_lake_at_rest_2048_stays_at_rest():
    """The headline bathymetry lowered by 4 m (every cell wet) with still water and walls: the
    well-balanced scheme keeps it at rest to FP32 round-off.  (A shoreline is not at rest in
    the reference's Boussinesq terms: the free-surface derivatives see the dry land's w = B.)"""
    sc = S.headline()
    bnd = [Boundary(SOLID, 2) for _ in range(4)]
    bottom = (sc.bottom - np.float32(4.0)).astype(np.float32)
    state = initial_state(sc.s, bottom, [])
    e = ShallowWaterEngine(sc.s, bnd, bottom, state)
    e.timestep(200)
    a = e.get_state()
    e.close()
    assert np.abs(a[1:]).max() < 2e-5
    assert np.abs(a[0] - state[0]).max() < 2e-6


def test_headline_full_10000_steps():
    """The headline case for its full 10 000 steps (the oracle cannot follow that far): the
    state stays finite and bounded, the waves reach the beach, and stepping in captured 50-step
    graphs gives the same bits as direct launches."""
    sc = S.headline()
    runs = []
    for graph_steps in (0, 50):
        e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        e.set_graph_steps(graph_steps)
        e.timestep(10000)
        runs.append(e.get_state())
        e.close()
    a = runs[0]
    assert np.isfinite(a).all()
    assert np.array_equal(a, runs[1])
    B = sc.bottom[2]
    deep = (sc.s.still_water_elevation - B) > 1.0
    eta_deep = np.where(deep, a[0] - sc.s.still_water_elevation, 0.0)
    assert np.abs(eta_deep).max() < 0.5                # Hs 0.1 m waves (0.09 m measured)
    assert np.abs(a[1][2:-2, 2:-2]).max() > 1e-3       # momentum has built up across the domain
    # the only water on land above +1 m is the millimetre film the reference's scheme keeps
    # where the east sponge meets a wall on the dry beach (about 150 cells)
    land_wet = ((a[0] - B) > 1e-4) & (B > 1.0)
    assert land_wet.sum() < 1000
