"""Shared helpers for comparing the HIP engine with the CPU oracle."""
import numpy as np

MIN_DEPTH = 1e-4          # wet <=> h = w - BA > 1e-4 m (compute.hlsl:1016)


def rel_err(a, b):
    """max |a - b| / max |b| per field (w as eta = w - SWL), over the ghost-inclusive planes."""
    out = {}
    for k, name in enumerate(("eta", "hu", "hv")):
        x, y = a[k].astype(np.float64), b[k].astype(np.float64)
        scale = max(np.abs(y).max(), 1e-30)
        out[name] = float(np.abs(x - y).max() / scale)
    return out


def wet_mask(state, bottom):
    return (state[0] - bottom[2]) > MIN_DEPTH


def eta(state, swl=0.0):
    s = state.copy()
    s[0] = s[0] - swl
    return s
