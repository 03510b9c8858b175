"""GetStats diagnostics and resetTimestep's adaptive dt (compute.hlsl:1319-1378,
engine.cpp:1536-1756).

CPU tests pin the oracle's reductions with closed forms (uniform states, the reference's
4x4-block domain) and invariants (mass of a closed basin); GPU tests compare the device's
FP64 reductions and the adaptive dt sequence with the oracle.
"""
import numpy as np
import pytest

from celeris_amd import SOLID, SPONGE, SINEWAVE, Boundary, settings
from celeris_amd import scenarios as S
from oracle.binding import Oracle

G = float(np.float32(9.81))   # gravity enters as FP32


def uniform_case(nx, ny, h0, q, r=0.0, dx=0.5):
    """Flat bottom at -h0, still surface, uniform momentum (q, r)."""
    W, L = dx * (nx - 1), dx * (ny - 1)
    s = settings(nx, ny, W, L, 0.01)
    X, Y = nx + 4, ny + 4
    bottom = np.full((3, Y, X), -h0, np.float32)
    state = np.zeros((3, Y, X), np.float32)
    state[1] = q
    state[2] = r
    return s, [Boundary(SOLID, 2) for _ in range(4)], bottom, state


def test_stats_uniform_closed_form():
    """h = h0 everywhere: sums are count * per-cell value, u = q/h0 (h0^4 >> epsilon), and the
    CFL rate is (|u| + sqrt(g h0)) / dx."""
    nx, ny, h0, q, r = 16, 12, 0.75, 0.3, -0.15
    s, bnd, bottom, state = uniform_case(nx, ny, h0, q, r)
    q, r = float(np.float32(q)), float(np.float32(r))      # the state enters as FP32
    st = Oracle(s, bnd, bottom, state, "f64").stats()
    n = nx * ny
    dx = 0.5
    u, v = q / h0, r / h0
    np.testing.assert_allclose(st[0], n * h0, rtol=1e-14)
    np.testing.assert_allclose(st[1], n * h0 * (-h0 + 0.5 * h0), rtol=1e-14)
    np.testing.assert_allclose(st[2], n * q, rtol=1e-14)
    np.testing.assert_allclose(st[3], n * r, rtol=1e-14)
    np.testing.assert_allclose(st[4], n * (q * u + r * v), rtol=1e-12)
    np.testing.assert_allclose(st[5], u * u + v * v, rtol=1e-12)
    np.testing.assert_allclose(st[6], h0, rtol=0)
    np.testing.assert_allclose(st[7], max(abs(u), abs(v)) / dx + np.sqrt(G * h0) / dx, rtol=1e-12)
    np.testing.assert_allclose(st[8], (u * u + v * v) / h0, rtol=1e-12)
    # the FP32 oracle forms the same per-cell values in float
    st32 = Oracle(s, bnd, bottom, state, "f32").stats()
    np.testing.assert_allclose(st32, st, rtol=2e-6)


def test_stats_domain_is_whole_4x4_blocks():
    """nx=10, ny=7: the reference's blocks cover columns 2..9 and rows 2..5 only."""
    nx, ny = 10, 7
    s, bnd, bottom, state = uniform_case(nx, ny, 1.0, 0.0)
    depth = np.zeros((ny + 4, nx + 4))
    rng = np.random.default_rng(1)
    depth[:] = rng.uniform(0.5, 1.5, depth.shape)
    bottom[:] = -depth.astype(np.float32)
    st = Oracle(s, bnd, bottom, state, "f64").stats()
    h = depth.astype(np.float32).astype(np.float64)
    assert st[0] == pytest.approx(h[2:6, 2:10].sum(), rel=1e-14)
    assert st[6] == h[2:6, 2:10].max()


def test_stats_dry_cells_and_lake_at_rest():
    """Dry cells add nothing; a lake at rest has zero momentum sums and zero speeds."""
    sc = S.small(nx=24, ny=20, soliton=False)
    st = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64").stats()
    BA = sc.bottom[2].astype(np.float64)
    h = np.maximum(0.0, sc.state[0].astype(np.float64) - BA)[2:22, 2:26]
    assert st[0] == pytest.approx(h.sum(), rel=1e-12)
    assert st[2] == 0 and st[3] == 0 and st[4] == 0 and st[5] == 0 and st[8] == 0
    assert st[7] == pytest.approx(np.sqrt(G * h.max()) / sc.s.width * (sc.s.nx - 1), rel=1e-6)


def bowl(nx=48, ny=40, dx=0.5):
    """Closed lake with a dry rim and a hump of water: every face flux of w stays inside,
    so the volume is conserved to round-off."""
    from celeris_amd import bottom_from_corners, initial_state
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, Y = S.corner_coords(nx, ny, W, L)
    R = 0.5 * min(W, L)
    z = -1.0 + 1.5 * ((X - W / 2) ** 2 + (Y - L / 2) ** 2) / R ** 2 + 0.05 * np.sin(X / 2.0) * np.cos(Y / 3.0)
    s = settings(nx, ny, W, L, S.cfl_dt(dx, 1.2))
    bottom = np.zeros((3, ny + 4, nx + 4), np.float32)
    from oracle.binding import bottom_from_corners as orc_bottom
    bottom = orc_bottom(nx, ny, z.astype(np.float32))
    state = np.zeros((3, ny + 4, nx + 4), np.float32)
    state[0] = np.maximum(0.0, bottom[2])
    xc = dx * (np.arange(nx + 4) - 2)
    yc = dx * (np.arange(ny + 4) - 2)
    hump = 0.06 * np.exp(-((xc[None, :] - 0.4 * W) ** 2 + (yc[:, None] - 0.5 * L) ** 2) / 4.0)
    wet = state[0] > bottom[2]
    state[0] = np.where(wet, state[0] + hump, state[0]).astype(np.float32)
    return s, [Boundary(SOLID, 2) for _ in range(4)], bottom, state


def test_stats_mass_conserved_in_closed_basin():
    s, bnd, bottom, state = bowl()
    o = Oracle(s, bnd, bottom, state, "f64")
    m0 = o.stats()[0]
    o.step(300)
    # nx, ny multiples of 4: the block domain is the whole interior.  sum(w - BA) is
    # conserved to 1e-12; sum(max(0, w - BA)) also clips round-off-negative depths
    assert abs(o.stats()[0] - m0) <= 1e-8 * m0
    assert o.stats()[2] != 0.0


def expected_dt(prev_old, target_cfl, rate, dt0, alpha=0.5, min_ratio=0.05):
    f = np.float32
    target = max(f(target_cfl) / f(rate), f(dt0) * f(min_ratio))
    if target > f(prev_old):
        return f(alpha) * target + (f(1) - f(alpha)) * f(prev_old)
    return target


def test_adaptive_dt_rule_in_oracle():
    """Euler bootstrap steps keep the uniform dt; afterwards dt follows the target-CFL rule
    with the moving average on growth (engine.cpp:1698-1726)."""
    sc = S.small(nx=32, ny=24, seed=5, time_scheme=3)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    dt0 = np.float32(sc.s.timestep)
    dt, st = o.update_timestep()
    assert dt == dt0 and o.scheme == 0
    target = np.float32(dt0) * np.float32(st[7])
    o.step(1)
    dt, _ = o.update_timestep()
    assert dt == dt0
    o.step(1)
    old = dt0
    for _ in range(12):
        assert o.scheme == 3
        dt, st = o.update_timestep()
        assert dt == expected_dt(old, target, np.float32(st[7]), dt0)
        old = dt
        o.step(1)
    assert np.isfinite(o.state()).all()


def test_uniform_dt_follows_set_dt():
    sc = S.small(nx=24, ny=24, seed=6)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    o.set_dt(0.5 * sc.s.timestep)
    o.step(3)
    dt, _ = o.update_timestep()
    assert dt == np.float32(0.5 * sc.s.timestep)


# ------------------------------------------------------------------ GPU

def _gpu_engine(sc):
    from celeris_amd import ShallowWaterEngine
    return ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(48, 40), (61, 37), (200, 130)])
def test_gpu_stats_match_oracle(nx, ny):
    bnd = [Boundary(SINEWAVE, 2, amplitude=0.03, period=1.5), Boundary(SPONGE, 8), Boundary(SOLID, 2),
           Boundary(SOLID, 2)]
    sc = S.small(nx=nx, ny=ny, seed=11, bnd=bnd)
    eng = _gpu_engine(sc)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    for steps in (0, 40, 80):
        eng.timestep(steps)
        o.step(steps)
        g = eng.stats()
        r = o.stats()
        names = ["sum_h", "sum_bh_h2", "sum_hu", "sum_hv", "sum_hu2v2", "max_u2v2", "max_h", "max_cfl_rate",
                 "max_u2v2_h"]
        gv = np.array([g[k] for k in names])
        scale = np.maximum(np.abs(r), 1e-30)
        # sums: same FP32 cell values, FP64 sums in another order; maxima: order-free
        assert np.all(np.abs(gv[:5] - r[:5]) <= 1e-5 * np.maximum(scale[:5], np.abs(r[0]) * 1e-3)), (steps, gv, r)
        assert np.all(np.abs(gv[5:] - r[5:]) <= 1e-5 * scale[5:]), (steps, gv, r)
        assert g["cells"] == 16 * (nx // 4) * (ny // 4)
        assert g["mass"] == pytest.approx(r[0] * 1000 * sc.s.width / (nx - 1) * sc.s.length / (ny - 1), rel=1e-5)
    eng.close()


@pytest.mark.gpu
def test_gpu_stats_bit_exact_state_gives_equal_sums():
    """Without trigonometric forcing the states are bit-identical, so only the FP64
    summation order differs: 1e-12 relative."""
    sc = S.small(nx=64, ny=48, seed=12)
    eng = _gpu_engine(sc)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    eng.timestep(60)
    o.step(60)
    g, r = eng.stats(), o.stats()
    assert g["sum_h"] == pytest.approx(r[0], rel=1e-12)
    assert g["sum_bh_h2"] == pytest.approx(r[1], rel=1e-12)
    assert g["max_cfl_rate"] == r[7] and g["max_h"] == r[6] and g["max_u2v2"] == r[5]
    eng.close()


@pytest.mark.gpu
def test_gpu_adaptive_dt_sequence_matches_oracle():
    """resetTimestep every frame on a PREDICTOR_ADAPTIVE run: the same dt sequence as the
    oracle and a bit-identical state (no trigonometric forcing)."""
    sc = S.small(nx=40, ny=36, seed=13, time_scheme=3)
    eng = _gpu_engine(sc)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    dts = []
    for _ in range(80):
        g = eng.update_timestep()
        dt, _ = o.update_timestep()
        assert np.float32(g["timestep"]) == dt
        dts.append(dt)
        eng.timestep(1)
        o.step(1)
    assert len(set(dts)) > 3, "dt never adapted"
    a, b = eng.get_state(), o.state()
    assert np.array_equal(a, b)
    assert eng.time == o.time
    eng.close()
