"""The two places where the workloads depart from the configurations' literal wording, each
checked on the FP32 oracle (the reference's arithmetic), so the reason is on record:

* the headline with sponges on the north and south sides as well: a sponge that crosses
  the dry beach divides 0 by 0 in the reference's damping factor at its outermost ghost
  row (compute.hlsl:1011-1049), and the state turns non-finite within a few steps; with
  walls there (the workload's choice) it stays finite;
* CFL 0.5: the reference's AB3 predictor is not stable there on the plain beach, which
  blows up within 1000 steps; at CFL 0.1 (the workloads' choice) it stays bounded.
"""
import numpy as np

from celeris_amd import scenarios as S
from oracle.binding import Oracle


def test_north_south_sponges_across_dry_land_go_non_finite():
    lit = S.headline(nx=256, ny=256, ns_sponge=True)
    o = Oracle(lit.s, lit.boundaries, lit.bottom, lit.state, "f32")
    o.step(3)
    assert not np.isfinite(o.state()).all()
    walls = S.headline(nx=256, ny=256)
    o = Oracle(walls.s, walls.boundaries, walls.bottom, walls.state, "f32")
    o.step(50)
    assert np.isfinite(o.state()).all()


def test_cfl_half_diverges_on_the_plain_beach():
    steep = S.plain_beach(cfl=0.5)
    o = Oracle(steep.s, steep.boundaries, steep.bottom, steep.state, "f32")
    o.step(1000)
    st = o.state()
    assert not np.isfinite(st).all() or np.abs(st[0]).max() > 1e3
    calm = S.plain_beach()
    o = Oracle(calm.s, calm.boundaries, calm.bottom, calm.state, "f32")
    o.step(1000)
    st = o.state()
    assert np.isfinite(st).all() and np.abs(st[0]).max() < 1.0
