"""cel_stats_noalloc: GetStats over caller-owned device planes with its partial sums in the
library's file-scope device array (no allocation in the call), against cel_engine_stats of
an engine holding the same state: the same kernels and reduction orders, so every field
must be identical.  Two calls in a row, the second on a smaller grid (the array keeps the
first call's partials), and the size limit."""
import ctypes

import numpy as np
import pytest

import celeris_amd as C
from celeris_amd import ShallowWaterEngine
from celeris_amd import scenarios as S

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime the library is linked against (already loaded by it), for the test's
    own device buffers."""
    C.load_library()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


H2D, D2H = 1, 2


def stats_noalloc(s, state, bottom):
    hip = hip_runtime()
    lib = C.load_library()
    planes = [np.ascontiguousarray(a, np.float32) for a in (state[0], state[1], state[2], bottom[2])]
    bufs = []
    try:
        for a in planes:
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
            bufs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        sums = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(sums), 9 * 8) == 0
        bufs.append(sums)
        C._check(lib.cel_stats_noalloc(ctypes.byref(s), *bufs[:4], s.nx + 4, sums, None))
        raw = np.zeros(9, np.float64)
        assert hip.hipMemcpy(raw.ctypes.data, sums, raw.nbytes, D2H) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
    out = C.CelStats()
    C._check(lib.cel_stats_from_sums(ctypes.byref(s), s.timestep, raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(out)))
    return out.as_dict()


def test_stats_noalloc_equals_engine_stats_twice():
    for nx, ny, seed in ((700, 530, 1), (96, 72, 2)):        # the second call is smaller
        sc = S.small(nx=nx, ny=ny, seed=seed)
        e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        e.timestep(40)
        ref = e.stats()
        st = e.get_state()
        e.close()
        got = stats_noalloc(sc.s, st, sc.bottom)
        assert ref["sum_hu"] != 0.0 and ref["max_cfl_rate"] > 0.0
        for k, v in ref.items():
            assert got[k] == v, (nx, ny, k, got[k], v)


def test_stats_noalloc_rejects_grids_above_its_scratch():
    lib = C.load_library()
    s = C.settings(4100, 4100, 100.0, 100.0, 0.01)
    sums = ctypes.c_void_p(16)
    dummy = ctypes.c_void_p(16)
    rc = lib.cel_stats_noalloc(ctypes.byref(s), dummy, dummy, dummy, dummy, 4104, sums, None)
    assert rc == 1 and b"4096" in lib.cel_last_error()
