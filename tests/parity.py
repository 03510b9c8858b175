"""Shared helpers for comparing the HIP engine with the CPU oracle."""
import numpy as np

MIN_DEPTH = 1e-4          # wet <=> h = w - BA > 1e-4 m (compute.hlsl:1016)


def rel_err(a, b):
    """Relative parity error per field over the ghost-inclusive planes: max |a - b| of eta
    (w - SWL) divided by max |eta|, and of hu and of hv divided by the largest momentum
    component max(|hu|, |hv|) (a field that stays ~0, such as hv under normal incidence,
    is measured against the flow's momentum scale rather than its own round-off)."""
    y = b.astype(np.float64)
    x = a.astype(np.float64)
    mom = max(np.abs(y[1]).max(), np.abs(y[2]).max(), 1e-30)
    scales = (max(np.abs(y[0]).max(), 1e-30), mom, mom)
    return {name: float(np.abs(x[k] - y[k]).max() / scales[k]) for k, name in enumerate(("eta", "hu", "hv"))}


def wet_mask(state, bottom):
    return (state[0] - bottom[2]) > MIN_DEPTH


def eta(state, swl=0.0):
    s = state.copy()
    s[0] = s[0] - swl
    return s
