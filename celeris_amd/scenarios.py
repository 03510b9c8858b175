"""Seeded synthetic scenarios (bathymetry at cell corners, forcing tables, initial state).

Each builder returns a Scenario: cel settings, four Boundary objects (W, E, S, N), the
bottom planes (BY, BX, BA) produced by the library's terrain preprocessing, and the
initial state (w, hu, hv) from the library's initial-condition routine.  Grids are
vertex-centred like the reference: dx = W/(nx-1), corner (i, j) at
x = (i-2.5) dx - W/2, y = (j-2.5) dy (terrain_heightfield.cpp:260-269).
"""
from dataclasses import dataclass, field

import numpy as np

from . import (EAST, IRREGULARWAVES, NORTH, SINEWAVE, SOLID, SOUTH, SPONGE, STATE_COMPENSATED, STATE_FP32, WEST,
               Boundary, bottom_from_corners, initial_state, settings)

G = 9.81


@dataclass
class Scenario:
    name: str
    s: object
    boundaries: list
    bottom: np.ndarray
    state: np.ndarray
    steps: int
    info: dict = field(default_factory=dict)


def corner_coords(nx, ny, W, L):
    """x', y' of the corners measured from the west/south edge cell centre (x' = x + W/2)."""
    i = np.arange(nx + 5, dtype=np.float64)
    j = np.arange(ny + 5, dtype=np.float64)
    xc = (i - 2.5) / (nx - 1) * W
    yc = (j - 2.5) / (ny - 1) * L
    return np.meshgrid(xc, yc)


def gaussian_random_field(shape, corr_cells, std, rng):
    """Zero-mean stationary field with a Gaussian covariance of the given correlation length."""
    ny, nx = shape
    noise = rng.standard_normal(shape)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    filt = np.exp(-2.0 * (np.pi * corr_cells) ** 2 * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    f -= f.mean()
    return f * (std / f.std())


def jonswap_waves(n, hs, tp, gamma, rng, dir_spread_deg=30.0, fmin_ratio=0.5, fmax_ratio=3.0):
    """n sinusoids from a JONSWAP spectrum: rows (amplitude, period, theta, phase), the
    irregular-wave table format of engine.cpp:3795-3798."""
    fp = 1.0 / tp
    f = np.linspace(fmin_ratio * fp, fmax_ratio * fp, n)
    df = f[1] - f[0]
    sigma = np.where(f <= fp, 0.07, 0.09)
    r = np.exp(-((f - fp) ** 2) / (2 * sigma ** 2 * fp ** 2))
    S = G ** 2 * (2 * np.pi) ** -4 * f ** -5 * np.exp(-1.25 * (fp / f) ** 4) * gamma ** r
    a = np.sqrt(2 * S * df)
    a *= (hs / 4.0) / np.sqrt(np.sum(a ** 2) / 2.0)        # Hs = 4 sqrt(m0), m0 = sum a^2/2
    theta = np.deg2rad(rng.uniform(-dir_spread_deg, dir_spread_deg, n))
    phase = rng.uniform(0, 2 * np.pi, n)
    return np.stack([a, 1.0 / f, theta, phase], axis=1).astype(np.float32)


def _finish(name, s, bnd, z, steps, solitons=(), info=None):
    bottom = bottom_from_corners(s.nx, s.ny, z.astype(np.float32))
    state = initial_state(s, bottom, solitons)
    return Scenario(name, s, bnd, bottom, state, steps, info or {})


# The AB3 predictor of the reference scheme is stable only for small Courant numbers: on a
# flat 1 m basin with a soliton it stays bounded at CFL 0.12 and diverges at 0.15 (checked
# with the CPU oracle).  The reference's own example cases run at CFL ~0.09
# (e.g. conical island: dx 0.1 m, dt 0.005 s, h 0.32 m), so the scenarios use 0.1.
DEFAULT_CFL = 0.1


def cfl_dt(dx, hmax, cfl=DEFAULT_CFL):
    return float(np.float32(cfl * dx / np.sqrt(G * hmax)))


def plain_beach(nx=256, ny=256, dx=0.5, steps=2000, cfl=DEFAULT_CFL):
    """Flat 1.0 m depth for x < 60 % of the domain, then a 1:20 slope up to a 0.3 m dry beach
    plateau; west sine wave (a 0.05 m, T 2 s), north/south walls, 20-cell east sponge,
    Bcoef 1/15, minmod theta 1.3, dt from the CFL number (0.1) on the 1 m depth."""
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, _ = corner_coords(nx, ny, W, L)
    z = np.where(X < 0.6 * W, -1.0, np.minimum(0.3, -1.0 + (X - 0.6 * W) / 20.0))
    s = settings(nx, ny, W, L, cfl_dt(dx, 1.0, cfl), theta=1.3)
    bnd = [Boundary(SINEWAVE, 2, 0.0, amplitude=0.05, period=2.0, theta=0.0),
           Boundary(SPONGE, 20), Boundary(SOLID, 2), Boundary(SOLID, 2)]
    return _finish("plain_beach", s, bnd, z, steps)


def headline(nx=2048, ny=2048, dx=1.0, steps=10000, seed=2048, n_waves=100, ns_sponge=False):
    """1:30 plane from 8 m depth to 2 m dry land at the east edge plus a seeded Gaussian
    random field (std 0.3 m, correlation 20 cells); west irregular waves from a JONSWAP
    spectrum (Hs 0.1 m, Tp 3 s, gamma 3.3, directions within +-30 deg); 20-cell east sponge.

    North and south are solid walls by default.  A south (or west) sponge that crosses
    dry land divides 0 by 0 in the reference's damping factor at its outermost ghost row
    (SpongeFunction(L, 0) = 0 and, with d = MIN_DEPTH, SpongeFunction(L, sqrt(g d) dt)
    rounds to 0 in FP32; compute.hlsl:1011-1049), so the reference itself turns that row
    into NaN within one step; ns_sponge=True builds that variant."""
    rng = np.random.default_rng(seed)
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, _ = corner_coords(nx, ny, W, L)
    z = np.maximum(-8.0, 2.0 - (W - X) / 30.0) + gaussian_random_field(X.shape, 20.0, 0.3, rng)
    waves = jonswap_waves(n_waves, 0.1, 3.0, 3.3, rng)
    hmax = float(-z.min())
    s = settings(nx, ny, W, L, cfl_dt(dx, hmax), theta=1.3)
    ns = Boundary(SPONGE, 20) if ns_sponge else Boundary(SOLID, 2)
    bnd = [Boundary(IRREGULARWAVES, 2, 0.0, waves=waves), Boundary(SPONGE, 20), ns, ns]
    return _finish("headline", s, bnd, z, steps, info={"seed": seed, "n_waves": n_waves, "hmax": hmax})


def solitary_channel(nx=4096, ny=1024, dx=0.25, steps=100000, H_over_h=0.3):
    """Channel: flat 1.0 m depth, then a 1:50 slope up to a 0.5 m dry plateau at the east end;
    a solitary wave (H/h = 0.3) starts near the west end moving east; solid walls all round."""
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, _ = corner_coords(nx, ny, W, L)
    z = np.minimum(0.5, np.maximum(-1.0, -1.0 + (X - (W - 75.0)) / 50.0))
    s = settings(nx, ny, W, L, cfl_dt(dx, 1.0 + H_over_h), theta=1.3)
    bnd = [Boundary(SOLID, 2) for _ in range(4)]
    sol = [(H_over_h * 1.0, 0.0, 1e6, 40.0 * dx * 4, 0.0)]   # (H, theta, crest length, xc, yc)
    return _finish("solitary_channel", s, bnd, z, steps, solitons=sol)


def small(nx=48, ny=40, dx=0.5, seed=0, bnd=None, soliton=True, steps=100, correction_steps=0,
          time_scheme=1, friction=0.0, is_manning=0, still_water=0.0, is_boussinesq=1):
    """Small seeded basin for parity tests: sloping, bumpy seabed that surfaces near the east
    edge, an optional oblique soliton, any boundary set."""
    rng = np.random.default_rng(seed)
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, Y = corner_coords(nx, ny, W, L)
    z = -0.8 + 1.0 * np.clip((X - 0.55 * W) / (0.45 * W), 0, None) ** 1.5
    z += 0.08 * np.sin(2 * np.pi * Y / (0.7 * L + 1.0)) + gaussian_random_field(X.shape, 3.0, 0.04, rng)
    s = settings(nx, ny, W, L, cfl_dt(dx, 1.0), theta=1.3, correction_steps_num=correction_steps,
                 time_scheme=time_scheme, friction=friction, is_manning=is_manning,
                 still_water_elevation=still_water, is_boussinesq=is_boussinesq)
    if bnd is None:
        bnd = [Boundary(SOLID, 2) for _ in range(4)]
    sol = [(0.1, 0.3, 1e6, 0.3 * W, 0.5 * L)] if soliton else []
    return _finish(f"small{nx}x{ny}", s, bnd, z, steps, solitons=sol)


def ensemble_member(rng, nx=None, ny=None, side_range=(128, 1024), dx=0.5, compensated=True, steps=5000,
                    name="member"):
    """One scenario of the batched ensemble: sides log-uniform in side_range (unless given), a
    seabed rising from h0 (1.5-3 m) to +0.5 m dry land at the east edge with a slope of 1:15
    to 1:60 (steepened where needed so that the beach spans at most 60 % of the domain and
    the west boundary sits at depth h0), plus a Gaussian random field (std 0.1-0.5 m,
    correlation 20 cells); west forcing is a sine wave or 32 JONSWAP components (amplitude
    0.02-0.2 m, period 1-6 s; the amplitude is capped at 3 % of the deep-water wavelength and
    5 % of h0 so that the non-breaking model stays in range); solid walls elsewhere;
    compensated state update."""
    lo, hi = np.log(side_range[0]), np.log(side_range[1])
    nx = nx or int(round(np.exp(rng.uniform(lo, hi))))
    ny = ny or int(round(np.exp(rng.uniform(lo, hi))))
    W, L = dx * (nx - 1), dx * (ny - 1)
    X, _ = corner_coords(nx, ny, W, L)
    h0 = rng.uniform(1.5, 3.0)
    slope = min(rng.uniform(15.0, 60.0), 0.6 * W / (h0 + 0.5))
    std = rng.uniform(0.1, 0.5)
    z = np.maximum(-h0, 0.5 - (W - X) / slope) + gaussian_random_field(X.shape, 20.0, std, rng)
    period = rng.uniform(1.0, 6.0)
    amp = min(rng.uniform(0.02, 0.2), 0.03 * G * period ** 2 / (2 * np.pi), 0.05 * h0)
    if rng.random() < 0.5:
        west = Boundary(SINEWAVE, 2, 0.0, amplitude=amp, period=period, theta=rng.uniform(-0.3, 0.3))
        forcing = "sine"
    else:
        west = Boundary(IRREGULARWAVES, 2, 0.0, waves=jonswap_waves(32, 2 * np.sqrt(2) * amp, period, 3.3, rng))
        forcing = "irregular"
    hmax = max(float(-z.min()), 0.5)
    s = settings(nx, ny, W, L, cfl_dt(dx, hmax), theta=1.3,
                 state_precision=STATE_COMPENSATED if compensated else STATE_FP32)
    bnd = [west, Boundary(SOLID, 2), Boundary(SOLID, 2), Boundary(SOLID, 2)]
    return _finish(name, s, bnd, z, steps, info={"slope": slope, "h0": h0, "grf_std": std, "amp": amp,
                                                 "period": period, "forcing": forcing})


def ensemble(n=64, seed=4, steps=5000, compensated=True):
    """The batched-ensemble workload: n independent members of uneven size."""
    rng = np.random.default_rng(seed)
    return [ensemble_member(rng, steps=steps, compensated=compensated, name=f"member{i}") for i in range(n)]


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    scenario: Scenario
    steps: int


def serving_requests(n, rate, seed=5, sides=(256, 512, 1024, 2048), weights=(0.4, 0.3, 0.2, 0.1),
                     steps_range=(500, 20000), compensated=True):
    """Scenario requests as a Poisson process of the given rate (1/s): square grids with the
    side drawn from `sides` with `weights`, durations log-uniform in steps_range, seabed and
    forcing drawn as for the ensemble."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        side = int(rng.choice(sides, p=np.asarray(weights) / np.sum(weights)))
        steps = int(round(np.exp(rng.uniform(np.log(steps_range[0]), np.log(steps_range[1])))))
        sc = ensemble_member(rng, nx=side, ny=side, steps=steps, compensated=compensated, name=f"req{i}")
        out.append(Request(i, t, sc, steps))
    return out


__all__ = ["ensemble_member", "ensemble", "Request", "serving_requests", "Scenario", "plain_beach", "headline", "solitary_channel", "small", "jonswap_waves",
           "gaussian_random_field", "corner_coords", "WEST", "EAST", "SOUTH", "NORTH"]
