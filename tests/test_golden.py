"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the FP32
oracle must reproduce them bit-for-bit on the CPU, and the HIP engine must reproduce them
on the GPU (bit-for-bit without trigonometric forcing, 1e-5 relative otherwise)."""
import glob
import os

import numpy as np
import pytest

from celeris_amd import Boundary, CelSettings
from oracle.binding import Oracle
from tests.parity import rel_err, wet_mask

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
TRIG = ("sine", "irregular")


def load(path):
    z = np.load(path, allow_pickle=False)
    s = CelSettings()
    for n, v in zip(z["settings_names"], z["settings_values"]):
        ftype = dict(CelSettings._fields_)[str(n)]
        setattr(s, str(n), int(v) if ftype.__name__ == "c_int" else float(v))
    bnd = []
    for i, row in enumerate(z["bnd"]):
        bnd.append(Boundary(int(row[0]), int(row[1]), float(row[2]), float(row[3]), float(row[4]), float(row[5]),
                            waves=z[f"waves{i}"] if f"waves{i}" in z.files else None,
                            uts=z[f"uts{i}"] if f"uts{i}" in z.files else None))
    return s, bnd, z["bottom"], z["state0"], z["state"], int(z["steps"]), float(z["time"])


def test_fixtures_present():
    assert len(FIXTURES) >= 7


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(path):
    s, bnd, bottom, st0, expect, steps, t = load(path)
    o = Oracle(s, bnd, bottom, st0, "f32")
    o.step(steps)
    assert np.array_equal(o.state(), expect)
    assert o.time == t


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_engine_reproduces_golden(path):
    from celeris_amd import ShallowWaterEngine
    s, bnd, bottom, st0, expect, steps, t = load(path)
    e = ShallowWaterEngine(s, bnd, bottom, st0)
    e.timestep(steps)
    got = e.get_state()
    assert e.time == t
    e.close()
    assert (wet_mask(got, bottom) == wet_mask(expect, bottom)).all()
    if any(k in os.path.basename(path) for k in TRIG):
        err = rel_err(got, expect)
        assert max(err.values()) <= 1e-5, err
    else:
        assert np.array_equal(got, expect)
