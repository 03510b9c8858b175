"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check or
time against; the product (celeris_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libceleris_oracle.so")
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


class OrcBoundary(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("width", ctypes.c_int), ("water_level", ctypes.c_float),
                ("amplitude", ctypes.c_float), ("period", ctypes.c_float), ("theta", ctypes.c_float),
                ("n_waves", ctypes.c_int), ("waves", _F), ("n_uts", ctypes.c_int), ("uts", _F)]


class OrcConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("width", ctypes.c_float), ("length", ctypes.c_float),
                ("still_water_elevation", ctypes.c_float), ("is_boussinesq", ctypes.c_int),
                ("epsilon", ctypes.c_float), ("theta", ctypes.c_float), ("friction", ctypes.c_float),
                ("is_manning", ctypes.c_int), ("correction_steps_num", ctypes.c_int), ("timestep", ctypes.c_float),
                ("time_scheme", ctypes.c_int), ("gravity", ctypes.c_float), ("bnd", OrcBoundary * 4)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for p in ("f32", "f64"):
            R = _F if p == "f32" else _D
            getattr(L, f"orc_{p}_create").restype = ctypes.c_void_p
            getattr(L, f"orc_{p}_create").argtypes = [ctypes.POINTER(OrcConfig), _F, _F]
            getattr(L, f"orc_{p}_destroy").argtypes = [ctypes.c_void_p]
            getattr(L, f"orc_{p}_step").argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, f"orc_{p}_set_dt").argtypes = [ctypes.c_void_p, ctypes.c_float]
            getattr(L, f"orc_{p}_get_state").argtypes = [ctypes.c_void_p, R]
            getattr(L, f"orc_{p}_time").argtypes = [ctypes.c_void_p]
            getattr(L, f"orc_{p}_time").restype = ctypes.c_double
            getattr(L, f"orc_{p}_scheme").argtypes = [ctypes.c_void_p]
            getattr(L, f"orc_{p}_coefficients").argtypes = [ctypes.c_void_p, R, R]
            getattr(L, f"orc_{p}_cr_solve").argtypes = [ctypes.c_int, R, R, R, R, R]
            getattr(L, f"orc_{p}_stats").argtypes = [ctypes.c_void_p, _D]
            getattr(L, f"orc_{p}_update_timestep").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 4 + [_D]
            getattr(L, f"orc_{p}_update_timestep").restype = ctypes.c_float
            getattr(L, f"orc_{p}_set_visual").argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
            getattr(L, f"orc_{p}_get_aux").argtypes = [ctypes.c_void_p, R]
            getattr(L, f"orc_{p}_get_normals").argtypes = [ctypes.c_void_p, R]
            getattr(L, f"orc_{p}_add_solitary_wave").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 4 + [ctypes.c_int]
            getattr(L, f"orc_{p}_refresh").argtypes = [ctypes.c_void_p]
            getattr(L, f"orc_{p}_set_bottom").argtypes = [ctypes.c_void_p, _F, ctypes.c_int]
            getattr(L, f"orc_{p}_raise_terrain").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 4 + [ctypes.c_int]
            getattr(L, f"orc_{p}_shift_terrain").argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int]
            getattr(L, f"orc_{p}_edit_water").argtypes = [ctypes.c_void_p] + [ctypes.c_float] * 5 + [ctypes.c_int]
            getattr(L, f"orc_{p}_get_bottom").argtypes = [ctypes.c_void_p, R]
            V = ctypes.c_float if p == "f32" else ctypes.c_double
            getattr(L, f"orc_{p}_ab3_increment").restype = V
            getattr(L, f"orc_{p}_ab3_increment").argtypes = [V] * 4
            getattr(L, f"orc_{p}_am4_increment").restype = V
            getattr(L, f"orc_{p}_am4_increment").argtypes = [V] * 5
            getattr(L, f"orc_{p}_second_order_derivatives").argtypes = [V] * 5 + [R]
            getattr(L, f"orc_{p}_time_series").argtypes = [_F, ctypes.c_int, ctypes.POINTER(ctypes.c_int), V, R]
            getattr(L, f"orc_{p}_irregular_strip").argtypes = [ctypes.c_void_p, ctypes.c_int, R]
        L.orc_bottom_from_corners.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_f32_minmod.restype = ctypes.c_float
        L.orc_f32_minmod.argtypes = [ctypes.c_float] * 3
        L.orc_f32_reconstruct.argtypes = [ctypes.c_float] * 4 + [_F, _F]
        L.orc_f32_correct_w.argtypes = [ctypes.c_float] * 3 + [_F, _F]
        L.orc_f32_divide_by_h.restype = ctypes.c_float
        L.orc_f32_divide_by_h.argtypes = [ctypes.c_float] * 2
        L.orc_f32_numerical_flux.restype = ctypes.c_float
        L.orc_f32_numerical_flux.argtypes = [ctypes.c_float] * 5
        L.orc_f32_sponge.restype = ctypes.c_float
        L.orc_f32_sponge.argtypes = [ctypes.c_float] * 2
        L.orc_f32_wavenumber.restype = ctypes.c_float
        L.orc_f32_wavenumber.argtypes = [ctypes.c_float] * 3
        L.orc_f32_ab_coef.argtypes = [ctypes.c_float] * 3 + [_F]
        _lib = L
    return _lib


def fptr(a):
    return a.ctypes.data_as(_F)


def dptr(a):
    return a.ctypes.data_as(_D)


class Oracle:
    """One simulation in the oracle, precision 'f32' (parity judge) or 'f64' (accuracy judge).

    boundaries: 4 objects with attributes type, width, water_level, amplitude, period,
    theta, waves (n x 4 float32 or None), uts (n x 4 float32 or None), indexed W, E, S, N.
    """

    def __init__(self, s, boundaries, bottom, state, precision="f32"):
        L = lib()
        self.p = precision
        self.nx, self.ny = s.nx, s.ny
        self._uniform_dt = s.timestep
        cfg = OrcConfig()
        for name, _ in OrcConfig._fields_[:-1]:
            setattr(cfg, name, getattr(s, name))
        self._keep = []
        for i, b in enumerate(boundaries):
            waves = None if b.waves is None else np.ascontiguousarray(b.waves, np.float32)
            uts = None if b.uts is None else np.ascontiguousarray(b.uts, np.float32)
            self._keep += [waves, uts]
            cfg.bnd[i] = OrcBoundary(b.type, b.width, b.water_level, b.amplitude, b.period, b.theta,
                                     0 if waves is None else waves.shape[0], None if waves is None else fptr(waves),
                                     0 if uts is None else uts.shape[0], None if uts is None else fptr(uts))
        bottom = np.ascontiguousarray(bottom, np.float32)
        state = np.ascontiguousarray(state, np.float32)
        self.h = getattr(L, f"orc_{precision}_create")(ctypes.byref(cfg), fptr(bottom), fptr(state))
        if not self.h:
            raise ValueError("oracle rejected the configuration")

    def __del__(self):
        if getattr(self, "h", None):
            getattr(lib(), f"orc_{self.p}_destroy")(self.h)
            self.h = None

    def step(self, n=1):
        getattr(lib(), f"orc_{self.p}_step")(self.h, n)

    def set_dt(self, dt):
        self._uniform_dt = dt
        getattr(lib(), f"orc_{self.p}_set_dt")(self.h, dt)

    def state(self):
        dt = np.float32 if self.p == "f32" else np.float64
        out = np.zeros((3, self.ny + 4, self.nx + 4), dt)
        ptr = fptr(out) if self.p == "f32" else dptr(out)
        getattr(lib(), f"orc_{self.p}_get_state")(self.h, ptr)
        return out

    @property
    def time(self):
        return getattr(lib(), f"orc_{self.p}_time")(self.h)

    @property
    def scheme(self):
        return getattr(lib(), f"orc_{self.p}_scheme")(self.h)

    def stats(self):
        """GetStats sums/maxima (9 doubles, order as in oracle_impl.inc)."""
        out = np.zeros(9, np.float64)
        getattr(lib(), f"orc_{self.p}_stats")(self.h, dptr(out))
        return out

    def update_timestep(self, ema_alpha=0.5, min_ratio=0.05, target_cfl=0.0, uniform_dt=None):
        """resetTimestep's dt choice; returns (dt, stats9)."""
        out = np.zeros(9, np.float64)
        u = self._uniform_dt if uniform_dt is None else uniform_dt
        dt = getattr(lib(), f"orc_{self.p}_update_timestep")(self.h, ema_alpha, min_ratio, target_cfl, u, dptr(out))
        return dt, out

    def set_visual(self, dissipation_threshold=0.25, whitewater_decay=0.1):
        getattr(lib(), f"orc_{self.p}_set_visual")(self.h, dissipation_threshold, whitewater_decay)

    def _planes(self, name):
        dt = np.float32 if self.p == "f32" else np.float64
        out = np.zeros((3, self.ny + 4, self.nx + 4), dt)
        getattr(lib(), f"orc_{self.p}_{name}")(self.h, fptr(out) if self.p == "f32" else dptr(out))
        return out

    def aux(self):
        """(max inundated depth, slope angle, whitewater) planes of the auxiliary texture."""
        return self._planes("get_aux")

    def normals(self):
        return self._planes("get_normals")

    def refresh(self):
        """Pass1 of the current state (H, U, V and the normal), as the bootstrap Pass1."""
        getattr(lib(), f"orc_{self.p}_refresh")(self.h)

    def set_bottom(self, bottom, keep_depth=True):
        bottom = np.ascontiguousarray(bottom, np.float32)
        getattr(lib(), f"orc_{self.p}_set_bottom")(self.h, fptr(bottom), int(keep_depth))

    # refresh=True re-runs Pass1 on the edited state, as the engine's next step sees it; the
    # reference's next step still uses the fluxes' inputs from before the edit (refresh=False)
    def raise_terrain(self, x, y, radius, displacement, refresh=False):
        getattr(lib(), f"orc_{self.p}_raise_terrain")(self.h, x, y, radius, displacement, int(refresh))

    def shift_terrain(self, shift_up, refresh=False):
        getattr(lib(), f"orc_{self.p}_shift_terrain")(self.h, shift_up, int(refresh))

    def edit_water(self, x, y, radius, disp_A, disp_B=0.0, refresh=False):
        getattr(lib(), f"orc_{self.p}_edit_water")(self.h, x, y, radius, disp_A, disp_B, int(refresh))

    def bottom(self):
        return self._planes("get_bottom")

    def add_solitary_wave(self, H, theta_deg, x0, y0, refresh=False):
        getattr(lib(), f"orc_{self.p}_add_solitary_wave")(self.h, H, theta_deg, x0, y0, int(refresh))

    def irregular_strip(self, side):
        """(w, hu, hv) planes one irregular-wave side writes at the current time (zero
        outside its strip)."""
        dt = np.float32 if self.p == "f32" else np.float64
        out = np.zeros((3, self.ny + 4, self.nx + 4), dt)
        getattr(lib(), f"orc_{self.p}_irregular_strip")(self.h, side, fptr(out) if self.p == "f32" else dptr(out))
        return out

    def coefficients(self):
        dt = np.float32 if self.p == "f32" else np.float64
        ax = np.zeros((3, self.ny + 4, self.nx + 4), dt)
        ay = np.zeros_like(ax)
        f = fptr if self.p == "f32" else dptr
        getattr(lib(), f"orc_{self.p}_coefficients")(self.h, f(ax), f(ay))
        return ax, ay


def cr_solve(a, b, c, d, precision="f32"):
    dt = np.float32 if precision == "f32" else np.float64
    a, b, c, d = (np.ascontiguousarray(v, dt) for v in (a, b, c, d))
    x = np.zeros_like(d)
    f = fptr if precision == "f32" else dptr
    getattr(lib(), f"orc_{precision}_cr_solve")(len(d), f(a), f(b), f(c), f(d), f(x))
    return x


def bottom_from_corners(nx, ny, z):
    z = np.ascontiguousarray(z, np.float32)
    out = np.zeros((3, ny + 4, nx + 4), np.float32)
    lib().orc_bottom_from_corners(nx, ny, fptr(z), fptr(out))
    return out


def set_threads(n):
    """OpenMP threads of the oracle's row loops (timing only; results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def ab3_increment(dt, f0, fm1, fm2, precision="f64"):
    return getattr(lib(), f"orc_{precision}_ab3_increment")(dt, f0, fm1, fm2)


def am4_increment(dt, f1, f0, fm1, fm2, precision="f64"):
    return getattr(lib(), f"orc_{precision}_am4_increment")(dt, f1, f0, fm1, fm2)


def second_order_derivatives(y_n, y_n1, y_n2, dt_n1, dt_n2, precision="f64"):
    dt = np.float32 if precision == "f32" else np.float64
    out = np.zeros(3, dt)
    getattr(lib(), f"orc_{precision}_second_order_derivatives")(y_n, y_n1, y_n2, dt_n1, dt_n2,
                                                              fptr(out) if precision == "f32" else dptr(out))
    return out


class TimeSeries:
    """The reference's uniform-time-series interpolation over rows (t, eta, hu, hv), with its
    forward-only cursor (starts at row 1)."""

    def __init__(self, rows, precision="f64"):
        self.rows = np.ascontiguousarray(rows, np.float32)
        self.p = precision
        self.iter = ctypes.c_int(1)

    def __call__(self, t):
        dt = np.float32 if self.p == "f32" else np.float64
        out = np.zeros(3, dt)
        getattr(lib(), f"orc_{self.p}_time_series")(fptr(self.rows), self.rows.shape[0], ctypes.byref(self.iter), t,
                                                    fptr(out) if self.p == "f32" else dptr(out))
        return out
