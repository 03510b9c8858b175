"""The C ABI library without a GPU: it loads, exports every function include/celeris.h
declares, and its host-side entry points (terrain preprocessing, initial condition, argument
checks) behave as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

import celeris_amd as C
from celeris_amd import scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "celeris.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cel_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.load_library()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(C.SIGNATURES), "ctypes signature table out of sync with the header"


def test_bad_arguments_report_einval():
    lib = C.load_library()
    out = np.zeros(10, np.float32)
    rc = lib.cel_bottom_from_corners(2, 2, C._fptr(out), C._fptr(out))
    assert rc == 1 and b"bad arguments" in lib.cel_last_error()
    rc = lib.cel_initial_state(None, C._fptr(out), None, 0, C._fptr(out))
    assert rc == 1
    with pytest.raises(C.CeleriError):
        C.bottom_from_corners(3, 3, np.zeros((8, 8), np.float32))


def test_null_engine_is_rejected():
    lib = C.load_library()
    assert lib.cel_engine_step(None, 1) == 1
    assert b"null engine" in lib.cel_last_error()


def test_engine_create_without_gpu_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    sc = S.small(nx=16, ny=16, soliton=False)
    with pytest.raises(C.CeleriError) as e:
        C.ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    assert e.value.status in (3, 2)


def test_initial_state_with_hot_layout():
    sc = S.small(nx=20, ny=18, soliton=True)
    st = sc.state
    assert st.shape == (3, 22, 24) and np.isfinite(st).all()
    assert st[1].max() > 0                     # soliton momentum
