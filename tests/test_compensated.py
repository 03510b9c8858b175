"""Compensated two-float state update (CEL_STATE_COMPENSATED).

The explicit update adds its increment to (w_hi, w_lo) with an exact two-sum, and the
line systems are solved for the increment of hu and hv (right-hand side without the
level-0 A, B, C terms), added the same way.  Checks: the variant stays within the FP32
parity bar of the FP32 oracle on ordinary cases; on a long-wave case over an elevated
datum, where per-step increments are a few ulps of the state, it tracks the FP64 oracle
more closely than the plain FP32 update; captured blocks reproduce it bit for bit.
"""
import numpy as np
import pytest

from celeris_amd import (SOLID, SPONGE, SINEWAVE, STATE_COMPENSATED, Boundary, CeleriError, ShallowWaterEngine,
                         settings)
from celeris_amd import scenarios as S
from oracle.binding import Oracle
from tests.parity import eta, rel_err, wet_mask


def comp(sc):
    sc.s.state_precision = STATE_COMPENSATED
    return sc


def test_compensated_rejects_corrector_steps():
    sc = comp(S.small(nx=24, ny=24, correction_steps=1))
    with pytest.raises(CeleriError, match="corrector"):
        ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)


def test_state_precision_validated():
    sc = S.small(nx=24, ny=24)
    sc.s.state_precision = 7
    with pytest.raises(CeleriError, match="state_precision"):
        ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)


def elevated_datum_case(nx=96, ny=16, datum=20.0, depth=10.0, amp=0.005, period=8.0):
    """Long, small sine wave over a flat bed 10 m below a 20 m datum: w ~ 20, per-step
    increments of w are O(10) ulps."""
    dx = 1.0
    W, L = dx * (nx - 1), dx * (ny - 1)
    s = settings(nx, ny, W, L, S.cfl_dt(dx, depth), still_water_elevation=datum)
    bottom = np.full((3, ny + 4, nx + 4), datum - depth, np.float32)
    state = np.zeros((3, ny + 4, nx + 4), np.float32)
    state[0] = datum
    bnd = [Boundary(SINEWAVE, 2, water_level=datum, amplitude=amp, period=period), Boundary(SPONGE, 20, water_level=datum),
           Boundary(SOLID, 2), Boundary(SOLID, 2)]
    return s, bnd, bottom, state


@pytest.mark.gpu
@pytest.mark.parametrize("seed,bnd", [(31, None),
                                      (32, [Boundary(SINEWAVE, 2, amplitude=0.03, period=1.5), Boundary(SPONGE, 8),
                                            Boundary(SOLID, 2), Boundary(SOLID, 2)])])
def test_compensated_within_fp32_bar(seed, bnd):
    sc = comp(S.small(nx=64, ny=48, seed=seed, bnd=bnd))
    eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    eng.timestep(400)
    o.step(400)
    a, b = eng.get_state(), o.state()
    err = rel_err(eta(a, 0.0), eta(b, 0.0))
    assert (wet_mask(a, sc.bottom) == wet_mask(b, sc.bottom)).all()
    assert max(err.values()) <= 1e-5, err
    eng.close()


@pytest.mark.gpu
def test_compensated_tracks_fp64_on_elevated_datum():
    s, bnd, bottom, state = elevated_datum_case()
    steps = 3000
    ref = Oracle(s, bnd, bottom, state, "f64")
    ref.step(steps)
    r = ref.state()
    errs = {}
    for prec in (0, STATE_COMPENSATED):
        s.state_precision = prec
        eng = ShallowWaterEngine(s, bnd, bottom, state)
        eng.timestep(steps)
        a = eng.get_state_f64()
        eng.close()
        inner = (slice(2, -2), slice(2, -22))     # outside the sponge
        e = a[0][inner] - r[0][inner]
        errs[prec] = (np.abs(e).max(), np.sqrt((e ** 2).mean()))
    amp = 0.005
    plain, compd = errs[0], errs[STATE_COMPENSATED]
    print("eta error / amplitude  plain FP32: Linf %.3e L2 %.3e   compensated: Linf %.3e L2 %.3e" %
          (plain[0] / amp, plain[1] / amp, compd[0] / amp, compd[1] / amp))
    assert compd[1] < 0.05 * plain[1]      # measured: 0.0075
    assert compd[0] < 1e-2 * amp


@pytest.mark.gpu
def test_compensated_graph_blocks_bit_identical():
    s, bnd, bottom, state = elevated_datum_case(nx=64, ny=24)
    s.state_precision = STATE_COMPENSATED
    outs = []
    for k in (0, 9):
        eng = ShallowWaterEngine(s, bnd, bottom, state)
        eng.set_graph_steps(k)
        eng.timestep(100)
        outs.append(eng.get_state_f64())
        eng.close()
    assert np.array_equal(outs[0], outs[1])
