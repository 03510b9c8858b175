"""The library's measurement probes: the HBM streaming kernels, the replay of the explicit
stencil's memory-access pattern and of the line solves' structure.  Their arguments are checked on the host (no GPU needed);
on the GPU each returns a positive launch time."""
import pytest

import celeris_amd as C


def test_pattern_probe_rejects_bad_arguments():
    lib = C.load_library()
    ms = C.ctypes.c_float()
    for args in ((2048, 2048, 10, 0), (2048, 2048, 12, 3), (3, 2048, 12, 0), (9000, 64, 12, 0)):
        assert lib.cel_probe_stencil_pattern(*args, 5, C.ctypes.byref(ms)) == 1, args
    assert lib.cel_probe_bandwidth(3, 1 << 20, 5, C.ctypes.byref(ms)) == 1
    for args in ((0, 1), (9000, 1), (16, 2), (16, -1)):
        assert lib.cel_probe_solve_pattern(*args, 5, C.ctypes.byref(ms)) == 1, args
    assert lib.cel_probe_solve_pattern(16, 1, 0, C.ctypes.byref(ms)) == 1


@pytest.mark.gpu
def test_probes_run():
    for kind in C.PROBE_KINDS:
        ms, gbs = C.probe_bandwidth(kind, 64 << 20, 3)
        assert ms > 0 and gbs > 0
    for layout in C.PATTERN_LAYOUTS:
        for rows in (8, 12, 16):
            assert C.probe_stencil_pattern(layout, 300, 200, rows, 3) > 0
    for eliminate in (False, True):
        assert C.probe_solve_pattern(64, eliminate, 3) > 0
