"""celeris_amd -- MI355X-native Celeris Boussinesq time step.

Python view of the C ABI in include/celeris.h (ctypes).  The numerical path runs in
libceleris_amd.so (hand-written HIP kernels for gfx950); there is no Python or CPU
fallback: importing the engine without the built library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CELERIS_AMD_LIB", os.path.join(HERE, "libceleris_amd.so"))

# BoundaryType (settings.hpp:90-96), BoundarySide, TimeIntegrationScheme (settings.hpp:304)
SOLID, SPONGE, SINEWAVE, IRREGULARWAVES, UNIFORMTIMESERIES = range(5)
WEST, EAST, SOUTH, NORTH = range(4)
EULER, PREDICTOR, CORRECTOR, PREDICTOR_ADAPTIVE = range(4)
STATE_FP32, STATE_COMPENSATED, STATE_FP64 = range(3)
STATUS = {0: "CEL_OK", 1: "CEL_EINVAL", 2: "CEL_ENOMEM", 3: "CEL_EDEVICE", 4: "CEL_ENUMERIC", 5: "CEL_ESTATE"}

_F = ctypes.POINTER(ctypes.c_float)


class CelBoundary(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("width", ctypes.c_int), ("water_level", ctypes.c_float),
                ("amplitude", ctypes.c_float), ("period", ctypes.c_float), ("theta", ctypes.c_float),
                ("n_waves", ctypes.c_int), ("waves", _F), ("n_uts", ctypes.c_int), ("uts", _F)]


class CelSettings(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("width", ctypes.c_float), ("length", ctypes.c_float),
                ("still_water_elevation", ctypes.c_float), ("is_boussinesq", ctypes.c_int),
                ("epsilon", ctypes.c_float), ("theta", ctypes.c_float), ("friction", ctypes.c_float),
                ("is_manning", ctypes.c_int), ("correction_steps_num", ctypes.c_int), ("timestep", ctypes.c_float),
                ("time_scheme", ctypes.c_int), ("gravity", ctypes.c_float), ("state_precision", ctypes.c_int)]


class CelSoliton(ctypes.Structure):
    _fields_ = [("param_h", ctypes.c_float), ("theta", ctypes.c_float), ("length", ctypes.c_float),
                ("xc", ctypes.c_float), ("yc", ctypes.c_float)]


class CelStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "sum_h", "sum_bh_h2", "sum_hu", "sum_hv", "sum_hu2v2", "max_u2v2", "max_h", "max_cfl_rate", "max_u2v2_h",
        "mass", "x_momentum", "y_momentum", "kinetic_energy", "potential_energy", "total_energy",
        "max_speed", "max_depth", "max_froude", "cfl_number")] + [("timestep", ctypes.c_float), ("cells", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CelAdaptive(ctypes.Structure):
    _fields_ = [("ema_alpha", ctypes.c_float), ("min_timestep_ratio", ctypes.c_float), ("target_cfl", ctypes.c_float)]


class CelTuning(ctypes.Structure):
    _fields_ = [("band_rows", ctypes.c_int), ("x_chunk", ctypes.c_int), ("x_waves", ctypes.c_int),
                ("y_chunk", ctypes.c_int), ("y_waves", ctypes.c_int), ("explicit_us", ctypes.c_float),
                ("x_lines_us", ctypes.c_float), ("y_lines_us", ctypes.c_float), ("fast_arith", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# Every symbol include/celeris.h declares, with its argument types.
SIGNATURES = {
    "cel_last_error": ([], ctypes.c_char_p),
    "cel_bottom_from_corners": ([ctypes.c_int, ctypes.c_int, _F, _F], ctypes.c_int),
    "cel_initial_state": ([ctypes.POINTER(CelSettings), _F, ctypes.POINTER(CelSoliton), ctypes.c_int, _F], ctypes.c_int),
    "cel_hot_start_state": ([ctypes.POINTER(CelSettings), _F, _F, ctypes.POINTER(CelSoliton), ctypes.c_int, _F],
                            ctypes.c_int),
    "cel_engine_sample": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _F], ctypes.c_int),
    "cel_engine_timestep_count": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_engine_set_visual": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "cel_engine_get_aux": ([ctypes.c_void_p, _F], ctypes.c_int),
    "cel_engine_get_normals": ([ctypes.c_void_p, _F], ctypes.c_int),
    "cel_engine_add_solitary_wave": ([ctypes.c_void_p] + [ctypes.c_float] * 4, ctypes.c_int),
    "cel_engine_set_bottom": ([ctypes.c_void_p, _F, ctypes.c_int], ctypes.c_int),
    "cel_engine_raise_terrain": ([ctypes.c_void_p] + [ctypes.c_float] * 4, ctypes.c_int),
    "cel_engine_shift_terrain": ([ctypes.c_void_p, ctypes.c_float], ctypes.c_int),
    "cel_engine_edit_water": ([ctypes.c_void_p] + [ctypes.c_float] * 5, ctypes.c_int),
    "cel_engine_create": ([ctypes.POINTER(CelSettings), ctypes.POINTER(CelBoundary), _F, _F,
                           ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "cel_engine_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "cel_engine_set_boundary": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CelBoundary)], ctypes.c_int),
    "cel_engine_set_dt": ([ctypes.c_void_p, ctypes.c_float], ctypes.c_int),
    "cel_engine_step": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_engine_sync": ([ctypes.c_void_p], ctypes.c_int),
    "cel_engine_get_state": ([ctypes.c_void_p, _F], ctypes.c_int),
    "cel_engine_set_state": ([ctypes.c_void_p, _F], ctypes.c_int),
    "cel_engine_get_state_f64": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cel_engine_time": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cel_engine_scheme": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_engine_timed_steps": ([ctypes.c_void_p, ctypes.c_int, _F], ctypes.c_int),
    "cel_engine_set_profiling": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_engine_kernel_time": ([ctypes.c_void_p, ctypes.c_char_p, _F, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_engine_device_bytes": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "cel_engine_coefficients": ([ctypes.c_void_p, _F, _F], ctypes.c_int),
    "cel_estimate_device_bytes": ([ctypes.POINTER(CelSettings), ctypes.POINTER(CelBoundary),
                                   ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "cel_engine_autotune": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CelTuning)], ctypes.c_int),
    "cel_engine_get_tuning": ([ctypes.c_void_p, ctypes.POINTER(CelTuning)], ctypes.c_int),
    "cel_engine_set_tuning": ([ctypes.c_void_p, ctypes.POINTER(CelTuning)], ctypes.c_int),
    "cel_batch_autotune": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_batch_create": ([ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "cel_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "cel_batch_add": ([ctypes.c_void_p, ctypes.POINTER(CelSettings), ctypes.POINTER(CelBoundary), _F, _F,
                       ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_batch_remove": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_batch_set_paused": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "cel_batch_step": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_batch_set_graph_steps": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_batch_sync": ([ctypes.c_void_p], ctypes.c_int),
    "cel_batch_timed_steps": ([ctypes.c_void_p, ctypes.c_int, _F], ctypes.c_int),
    "cel_batch_count": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_batch_device_bytes": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "cel_batch_get_state": ([ctypes.c_void_p, ctypes.c_int, _F], ctypes.c_int),
    "cel_batch_get_state_f64": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cel_batch_stats": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CelStats)], ctypes.c_int),
    "cel_batch_time": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "cel_engine_set_graph_steps": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cel_engine_stats": ([ctypes.c_void_p, ctypes.POINTER(CelStats)], ctypes.c_int),
    "cel_engine_reset_timestep": ([ctypes.c_void_p, ctypes.POINTER(CelAdaptive), ctypes.POINTER(CelStats)],
                                  ctypes.c_int),
    "cel_probe_bandwidth": ([ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _F], ctypes.c_int),
    "cel_probe_stencil_pattern": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _F],
                                  ctypes.c_int),
    "cel_probe_solve_pattern": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _F], ctypes.c_int),
    "cel_engine_check_numeric": ([ctypes.c_void_p], ctypes.c_int),
    "cel_batch_check_numeric": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_batch_set_state": ([ctypes.c_void_p, ctypes.c_int, _F], ctypes.c_int),
    "cel_batch_adopt": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cel_engine_checkpoint": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "cel_engine_restore": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t], ctypes.c_int),
    "cel_stats_noalloc": ([ctypes.POINTER(CelSettings), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "cel_stats_from_sums": ([ctypes.POINTER(CelSettings), ctypes.c_float, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(CelStats)], ctypes.c_int),
}

_lib = None


def load_library(path=LIB_PATH):
    """Loads libceleris_amd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m celeris_amd.build` "
                           "(the HIP library is the only implementation of the step)")
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class CeleriError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def _check(status):
    if status != 0:
        raise CeleriError(status, load_library().cel_last_error().decode())


def _fptr(a):
    return a.ctypes.data_as(_F)


def settings(nx, ny, width, length, timestep, still_water_elevation=0.0, is_boussinesq=1, epsilon=5e-8, theta=1.3,
             friction=0.0, is_manning=0, correction_steps_num=0, time_scheme=PREDICTOR, gravity=9.81,
             state_precision=STATE_FP32):
    """cel_settings with the reference's defaults (settings.hpp:376-435; g from gui_manager.cpp:107)."""
    return CelSettings(nx, ny, width, length, still_water_elevation, is_boussinesq, epsilon, theta, friction,
                       is_manning, correction_steps_num, timestep, time_scheme, gravity, state_precision)


class Boundary:
    """BoundarySetting (settings.hpp:149-159)."""

    def __init__(self, type=SOLID, width=2, water_level=0.0, amplitude=0.0, period=0.0, theta=0.0, waves=None,
                 uts=None):
        self.type, self.width, self.water_level = type, width, water_level
        self.amplitude, self.period, self.theta = amplitude, period, theta
        self.waves = None if waves is None else np.ascontiguousarray(waves, dtype=np.float32).reshape(-1, 4)
        self.uts = None if uts is None else np.ascontiguousarray(uts, dtype=np.float32).reshape(-1, 4)

    def to_c(self):
        nw = 0 if self.waves is None else self.waves.shape[0]
        nu = 0 if self.uts is None else self.uts.shape[0]
        return CelBoundary(self.type, self.width, self.water_level, self.amplitude, self.period, self.theta, nw,
                           None if nw == 0 else _fptr(self.waves), nu, None if nu == 0 else _fptr(self.uts))


def bottom_from_corners(nx, ny, z):
    """UpdateTerrainHeightfield (terrain_heightfield.cpp:244-298): corners -> (BY, BX, BA)."""
    z = np.ascontiguousarray(z, dtype=np.float32)
    assert z.shape == (ny + 5, nx + 5)
    out = np.zeros((3, ny + 4, nx + 4), np.float32)
    _check(load_library().cel_bottom_from_corners(nx, ny, _fptr(z), _fptr(out)))
    return out


def initial_state(s, bottom, solitons=()):
    """createSimTextures' initial condition (engine.cpp:2945-3016)."""
    bottom = np.ascontiguousarray(bottom, dtype=np.float32)
    out = np.zeros((3, s.ny + 4, s.nx + 4), np.float32)
    arr = (CelSoliton * max(1, len(solitons)))(*[CelSoliton(*sol) for sol in solitons])
    _check(load_library().cel_initial_state(ctypes.byref(s), _fptr(bottom), arr, len(solitons), _fptr(out)))
    return out


def hot_start_state(s, bottom, hot, solitons=()):
    """Initial condition from a hot-start state plus solitary waves (engine.cpp:2943-3016)."""
    bottom = np.ascontiguousarray(bottom, dtype=np.float32)
    hot = np.ascontiguousarray(hot, dtype=np.float32)
    out = np.zeros((3, s.ny + 4, s.nx + 4), np.float32)
    arr = (CelSoliton * max(1, len(solitons)))(*[CelSoliton(*sol) for sol in solitons])
    _check(load_library().cel_hot_start_state(ctypes.byref(s), _fptr(bottom), _fptr(hot), arr, len(solitons),
                                              _fptr(out)))
    return out


PROBE_KINDS = {"read": 0, "write": 1, "copy": 2}


def probe_bandwidth(kind, nbytes=1 << 30, launches=20):
    """Average launch time (ms) and rate (GB/s of bytes moved: read or written bytes, twice the
    size for a copy) of the library's HBM streaming probe over a device buffer of nbytes."""
    ms = ctypes.c_float()
    _check(load_library().cel_probe_bandwidth(PROBE_KINDS[kind], nbytes, launches, ctypes.byref(ms)))
    moved = nbytes * (2 if kind == "copy" else 1)
    return ms.value, moved / (ms.value * 1e-3) / 1e9


PATTERN_LAYOUTS = {"stencil": 0, "row_major_hv_dy": 1, "aligned_64": 2}


def probe_stencil_pattern(layout="stencil", nx=2048, ny=2048, band_rows=12, launches=20):
    """Average launch time (ms) of the explicit stencil's memory-access pattern without its
    arithmetic (cel_probe_stencil_pattern) in one of PATTERN_LAYOUTS."""
    ms = ctypes.c_float()
    _check(load_library().cel_probe_stencil_pattern(nx, ny, band_rows, PATTERN_LAYOUTS[layout], launches,
                                                    ctypes.byref(ms)))
    return ms.value


def probe_solve_pattern(lines=2048, eliminate=True, launches=20):
    """Average launch time (ms) of the line solves' structure without their coefficient
    arithmetic (cel_probe_solve_pattern): `lines` x-lines and as many y-lines of 2048 unknowns,
    equations combined by cyclic reduction's eliminations or by additions only."""
    ms = ctypes.c_float()
    _check(load_library().cel_probe_solve_pattern(lines, 1 if eliminate else 0, launches, ctypes.byref(ms)))
    return ms.value


class ShallowWaterEngine:
    """The step, settings and boundary interface of ShallowWaterEngine (engine.hpp:58-97) on one GPU."""

    def __init__(self, s, boundaries, bottom, state):
        lib = load_library()
        self.s = s
        self.nx, self.ny = s.nx, s.ny
        self._bnd = list(boundaries)
        cb = (CelBoundary * 4)(*[b.to_c() for b in self._bnd])
        bottom = np.ascontiguousarray(bottom, dtype=np.float32)
        state = np.ascontiguousarray(state, dtype=np.float32)
        h = ctypes.c_void_p()
        _check(lib.cel_engine_create(ctypes.byref(s), cb, _fptr(bottom), _fptr(state), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _check(load_library().cel_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_boundary(self, side, b):
        self._bnd[side] = b
        c = b.to_c()
        _check(load_library().cel_engine_set_boundary(self._h, side, ctypes.byref(c)))

    def reset_timestep(self, dt):
        _check(load_library().cel_engine_set_dt(self._h, dt))

    def timestep(self, n=1):
        _check(load_library().cel_engine_step(self._h, n))

    def sync(self):
        _check(load_library().cel_engine_sync(self._h))

    def get_state(self):
        out = np.zeros((3, self.ny + 4, self.nx + 4), np.float32)
        _check(load_library().cel_engine_get_state(self._h, _fptr(out)))
        return out

    def get_state_f64(self):
        """hi + lo words in FP64 (compensated engines), else the FP32 state widened."""
        out = np.zeros((3, self.ny + 4, self.nx + 4), np.float64)
        _check(load_library().cel_engine_get_state_f64(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.float32)
        _check(load_library().cel_engine_set_state(self._h, _fptr(state)))

    def sample(self, cells):
        """(w, hu, hv) at ghost-inclusive cells [(i, j), ...], read on the device: shape (n, 3)."""
        xy = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 2))
        out = np.zeros((len(xy), 3), np.float32)
        if len(xy):
            _check(load_library().cel_engine_sample(self._h, xy.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                     len(xy), _fptr(out)))
        return out

    @property
    def time(self):
        t = ctypes.c_double()
        _check(load_library().cel_engine_time(self._h, ctypes.byref(t)))
        return t.value

    @property
    def scheme(self):
        v = ctypes.c_int()
        _check(load_library().cel_engine_scheme(self._h, ctypes.byref(v)))
        return v.value

    @property
    def timestep_count(self):
        """The reference's timestep_count: 1 at creation, +1 per completed predictor cycle."""
        v = ctypes.c_int()
        _check(load_library().cel_engine_timestep_count(self._h, ctypes.byref(v)))
        return v.value

    def set_visual(self, enable=True, dissipation_threshold=0.25, whitewater_decay=0.1):
        """Turns the per-step visual outputs (inundation record, slope angle, whitewater) on/off."""
        _check(load_library().cel_engine_set_visual(self._h, int(enable), dissipation_threshold, whitewater_decay))

    def aux(self):
        """(max inundated depth, slope angle, whitewater) planes, ghost-inclusive."""
        out = np.zeros((3, self.ny + 4, self.nx + 4), np.float32)
        _check(load_library().cel_engine_get_aux(self._h, _fptr(out)))
        return out

    def normals(self):
        out = np.zeros((3, self.ny + 4, self.nx + 4), np.float32)
        _check(load_library().cel_engine_get_normals(self._h, _fptr(out)))
        return out

    def add_solitary_wave(self, H, theta_deg, x0, y0):
        _check(load_library().cel_engine_add_solitary_wave(self._h, H, theta_deg, x0, y0))

    def set_bottom(self, bottom, keep_depth=True):
        """New bottom planes (BY, BX, BA); keep_depth moves the water with the terrain."""
        bottom = np.ascontiguousarray(bottom, dtype=np.float32)
        _check(load_library().cel_engine_set_bottom(self._h, _fptr(bottom), int(keep_depth)))

    def raise_terrain(self, x, y, radius, displacement):
        _check(load_library().cel_engine_raise_terrain(self._h, x, y, radius, displacement))

    def shift_terrain(self, shift_up):
        _check(load_library().cel_engine_shift_terrain(self._h, shift_up))

    def edit_water(self, x, y, radius, disp_A, disp_B=0.0):
        """Left-mouse brush: add water (s, 0), remove (-s, 0), stir (-s, 1.5 s)."""
        _check(load_library().cel_engine_edit_water(self._h, x, y, radius, disp_A, disp_B))

    def timed_steps(self, n):
        ms = ctypes.c_float()
        _check(load_library().cel_engine_timed_steps(self._h, n, ctypes.byref(ms)))
        return ms.value

    def checkpoint(self):
        """Everything the next steps read, as a byte array (see cel_engine_checkpoint)."""
        n = ctypes.c_size_t()
        _check(load_library().cel_engine_checkpoint(self._h, None, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _check(load_library().cel_engine_checkpoint(self._h, buf.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        return buf

    def restore(self, buf):
        """Continues from a checkpoint of an engine with the same inputs, bit-identically."""
        buf = np.ascontiguousarray(buf, np.uint8)
        _check(load_library().cel_engine_restore(self._h, buf.ctypes.data_as(ctypes.c_void_p), buf.nbytes))

    def numeric_ok(self):
        """False once a step has produced a non-finite value (the device sentinel); waits for
        the engine's work."""
        rc = load_library().cel_engine_check_numeric(self._h)
        if rc == 4:
            return False
        _check(rc)
        return True

    def set_profiling(self, on):
        _check(load_library().cel_engine_set_profiling(self._h, int(bool(on))))

    def kernel_time(self, name):
        ms, n = ctypes.c_float(), ctypes.c_int()
        _check(load_library().cel_engine_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def device_bytes(self):
        b = ctypes.c_size_t()
        _check(load_library().cel_engine_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def autotune(self, repeats=10):
        """Times the band heights and line-solve block shapes on the current state and keeps
        the fastest (results are unchanged); returns the choice as a dict."""
        t = CelTuning()
        _check(load_library().cel_engine_autotune(self._h, repeats, ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        t = CelTuning()
        _check(load_library().cel_engine_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, band_rows, x_chunk, x_waves, y_chunk, y_waves, fast_arith=True):
        """Fixes the work decomposition; fast_arith=False forbids the explicit kernel's
        shortened exact division and square root (results are the same either way)."""
        t = CelTuning(band_rows, x_chunk, x_waves, y_chunk, y_waves, 0.0, 0.0, 0.0, 1 if fast_arith else 0)
        _check(load_library().cel_engine_set_tuning(self._h, ctypes.byref(t)))

    def set_graph_steps(self, k):
        """Run cel_engine_step in captured blocks of k steps (0: step by step)."""
        _check(load_library().cel_engine_set_graph_steps(self._h, k))

    def stats(self):
        """GetStats of the current state (FP64 device reductions) as a dict."""
        st = CelStats()
        _check(load_library().cel_engine_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def update_timestep(self, ema_alpha=0.5, min_timestep_ratio=0.05, target_cfl=0.0):
        """resetTimestep: GetStats, then the dt of the next step (adaptive or uniform)."""
        st = CelStats()
        a = CelAdaptive(ema_alpha, min_timestep_ratio, target_cfl)
        _check(load_library().cel_engine_reset_timestep(self._h, ctypes.byref(a), ctypes.byref(st)))
        return st.as_dict()

    def coefficients(self):
        ax = np.zeros((3, self.ny + 4, self.nx + 4), np.float32)
        ay = np.zeros_like(ax)
        _check(load_library().cel_engine_coefficients(self._h, _fptr(ax), _fptr(ay)))
        return ax, ay


def estimate_device_bytes(s, boundaries):
    """Device bytes an engine with these settings and boundaries holds once created."""
    cb = (CelBoundary * 4)(*[b.to_c() for b in boundaries])
    out = ctypes.c_size_t()
    _check(load_library().cel_estimate_device_bytes(ctypes.byref(s), cb, ctypes.byref(out)))
    return out.value


class Batch:
    """Independent scenarios advanced together (cel_batch_*): one launch per kernel per step
    over all scenarios, optionally replayed from captured multi-step graphs."""

    def __init__(self, graph_steps=0):
        h = ctypes.c_void_p()
        _check(load_library().cel_batch_create(ctypes.byref(h)))
        self._h = h
        self._dims = {}
        if graph_steps:
            self.set_graph_steps(graph_steps)

    def close(self):
        if getattr(self, "_h", None):
            _check(load_library().cel_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, s, boundaries, bottom, state):
        cb = (CelBoundary * 4)(*[b.to_c() for b in boundaries])
        bottom = np.ascontiguousarray(bottom, dtype=np.float32)
        state = np.ascontiguousarray(state, dtype=np.float32)
        i = ctypes.c_int()
        _check(load_library().cel_batch_add(self._h, ctypes.byref(s), cb, _fptr(bottom), _fptr(state), ctypes.byref(i)))
        self._dims[i.value] = (s.nx, s.ny)
        return i.value

    def adopt(self, engine):
        """Moves a ShallowWaterEngine (FP32 or compensated, created anywhere, e.g. on another
        thread while the batch runs) into the batch; returns its member id.  The engine
        object is closed."""
        i = ctypes.c_int()
        _check(load_library().cel_batch_adopt(self._h, engine._h, ctypes.byref(i)))
        self._dims[i.value] = (engine.nx, engine.ny)
        engine.close()
        return i.value

    def remove(self, sid):
        _check(load_library().cel_batch_remove(self._h, sid))
        self._dims.pop(sid, None)

    def set_paused(self, sid, paused=True):
        """A paused scenario keeps its state and is skipped by step() until resumed."""
        _check(load_library().cel_batch_set_paused(self._h, sid, int(paused)))

    def step(self, n=1):
        _check(load_library().cel_batch_step(self._h, n))

    def autotune(self, repeats=10, tune_lines=True):
        """Band height of the batched explicit launch (and each scenario's line-solve shapes)."""
        r = ctypes.c_int()
        _check(load_library().cel_batch_autotune(self._h, repeats, int(bool(tune_lines)), ctypes.byref(r)))
        return r.value

    def set_graph_steps(self, k):
        _check(load_library().cel_batch_set_graph_steps(self._h, k))

    def sync(self):
        _check(load_library().cel_batch_sync(self._h))

    def timed_steps(self, n):
        ms = ctypes.c_float()
        _check(load_library().cel_batch_timed_steps(self._h, n, ctypes.byref(ms)))
        return ms.value

    def __len__(self):
        n = ctypes.c_int()
        _check(load_library().cel_batch_count(self._h, ctypes.byref(n)))
        return n.value

    def device_bytes(self):
        b = ctypes.c_size_t()
        _check(load_library().cel_batch_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def _dim(self, sid):
        if sid not in self._dims:
            raise CeleriError(1, f"no such scenario {sid}")
        return self._dims[sid]

    def get_state(self, sid):
        nx, ny = self._dim(sid)
        out = np.zeros((3, ny + 4, nx + 4), np.float32)
        _check(load_library().cel_batch_get_state(self._h, sid, _fptr(out)))
        return out

    def set_state(self, sid, state):
        """Hot start of one member (clears its numeric sentinel)."""
        nx, ny = self._dim(sid)
        state = np.ascontiguousarray(state, dtype=np.float32)
        assert state.shape == (3, ny + 4, nx + 4)
        _check(load_library().cel_batch_set_state(self._h, sid, _fptr(state)))

    def failed(self):
        """Ids of the members whose steps produced a non-finite value (device sentinel); waits
        for the batch's work."""
        n = ctypes.c_int()
        cap = max(1, len(self._dims))
        ids = (ctypes.c_int * cap)()
        _check(load_library().cel_batch_check_numeric(self._h, ids, cap, ctypes.byref(n)))
        return [ids[i] for i in range(min(n.value, cap))]

    def get_state_f64(self, sid):
        nx, ny = self._dim(sid)
        out = np.zeros((3, ny + 4, nx + 4), np.float64)
        _check(load_library().cel_batch_get_state_f64(self._h, sid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def stats(self, sid):
        st = CelStats()
        _check(load_library().cel_batch_stats(self._h, sid, ctypes.byref(st)))
        return st.as_dict()

    def time(self, sid):
        t = ctypes.c_double()
        _check(load_library().cel_batch_time(self._h, sid, ctypes.byref(t)))
        return t.value
