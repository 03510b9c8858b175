"""Writes the golden fixtures: small seeded scenarios and the FP32 oracle's state after a
fixed number of steps (inputs and expected outputs in one compressed .npz per case).

    python tests/golden/make_golden.py

The fixtures pin the oracle against later edits (tests/test_golden.py) and give the GPU
tests expected outputs that do not need the oracle at run time.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from celeris_amd import IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, UNIFORMTIMESERIES, Boundary  # noqa: E402
from celeris_amd import scenarios as S  # noqa: E402
from oracle.binding import Oracle  # noqa: E402


def cases():
    t = np.arange(0, 20, 0.05, dtype=np.float32)
    uts = np.stack([t, 0.02 * np.sin(2 * np.pi * t / 2.0), 0.04 * np.sin(2 * np.pi * t / 2.0), 0 * t], 1)
    waves = S.jonswap_waves(9, 0.04, 1.8, 3.3, np.random.default_rng(12))
    yield "solid_32", dict(nx=32, ny=32, seed=1), 100
    yield "sponge_32", dict(nx=32, ny=32, seed=2, bnd=[Boundary(SPONGE, 6), Boundary(SPONGE, 6), Boundary(SOLID, 2),
                                                       Boundary(SPONGE, 5)]), 100
    yield "sine_32", dict(nx=32, ny=32, seed=3, soliton=False,
                          bnd=[Boundary(SINEWAVE, 2, amplitude=0.02, period=1.4, theta=0.3), Boundary(SPONGE, 8),
                               Boundary(SOLID, 2), Boundary(SOLID, 2)]), 100
    yield "irregular_32", dict(nx=32, ny=32, seed=4, soliton=False,
                               bnd=[Boundary(IRREGULARWAVES, 2, waves=waves), Boundary(SPONGE, 8), Boundary(SOLID, 2),
                                    Boundary(SOLID, 2)]), 100
    yield "uts_32", dict(nx=32, ny=32, seed=5, soliton=False,
                         bnd=[Boundary(UNIFORMTIMESERIES, 2, uts=uts), Boundary(SOLID, 2), Boundary(SOLID, 2),
                              Boundary(SOLID, 2)]), 100
    yield "corrector_32", dict(nx=32, ny=32, seed=6, correction_steps=2), 60
    yield "solid_64_10", dict(nx=64, ny=64, seed=7), 10


def bnd_arrays(bnd):
    rows = []
    for b in bnd:
        rows.append([b.type, b.width, b.water_level, b.amplitude, b.period, b.theta])
    return np.array(rows, np.float64)


def main():
    for name, kw, steps in cases():
        sc = S.small(**kw)
        o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
        o.step(steps)
        extra = {}
        for i, b in enumerate(sc.boundaries):
            if b.waves is not None:
                extra[f"waves{i}"] = b.waves
            if b.uts is not None:
                extra[f"uts{i}"] = b.uts
        settings = {n: getattr(sc.s, n) for n, _ in sc.s._fields_}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), bottom=sc.bottom, state0=sc.state, state=o.state(),
                            steps=steps, time=o.time, bnd=bnd_arrays(sc.boundaries),
                            settings_names=np.array(list(settings)), settings_values=np.array(list(settings.values()),
                                                                                                np.float64), **extra)
        print(name, steps)


if __name__ == "__main__":
    main()
