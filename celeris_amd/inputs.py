"""Reference input formats: CML experiment files, .cbf bathymetry, irregular-wave tables,
uniform-time-series tables and hot-start files, turned into the settings, boundaries,
bottom planes and initial state the engine takes.

Readers follow the reference's parsers:
  CML (main.cpp:500-913, tinyxml): elements with attributes; attribute values may be
      unquoted (`epsilon = 5e-8`), which tinyxml accepts and an XML parser does not, so a
      small tag reader is used here.
  .cbf (terrain_heightfield.cpp:136-209): `[tag] value` header lines up to a line that
      starts with '=', then nx*ny heights, row by row (x fastest).
  bathymetry sampling (getBathymetry, terrain_heightfield.cpp:213-243): bilinear in FP32
      with clamping at the edges, at the cell corners UpdateTerrainHeightfield visits.
  irregular-wave table (engine.cpp:3518-3543, 3777-3856): header with
      `[NumberOfWaves] n` (at most 1000) up to a '=' line, then rows of
      (amplitude, period, theta, phase).
  uniform time series (engine.cpp:3545-3556, 3858-3909): header up to a '=' line, then
      rows of (t, eta, hu, hv).
  hot start (engine.cpp:2943-2970): per ghost-inclusive cell, row by row:
      i j w hu hv alpha.
Files ending in .gz are read transparently.
"""
import gzip
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import (IRREGULARWAVES, SINEWAVE, SOLID, SPONGE, UNIFORMTIMESERIES, PREDICTOR, PREDICTOR_ADAPTIVE, Boundary,
               bottom_from_corners, hot_start_state, initial_state, settings)

PI_F = np.float32(math.atan(1.0) * 4.0)   # the reference's PI = atan(1.0f) * 4.0f, main.cpp:78
MAX_IRREGULAR_WAVES = 1000


def _open_text(path):
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        path = path + ".gz"
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


# ------------------------------------------------------------------ CML

@dataclass
class Element:
    name: str
    attrs: dict
    text: str = ""
    children: list = field(default_factory=list)

    def child(self, name):
        for c in self.children:
            if c.name == name:
                return c
        return None

    def all(self, name):
        return [c for c in self.children if c.name == name]


_TAG = re.compile(r"<(/?)([A-Za-z_][\w.-]*)((?:[^>\"']|\"[^\"]*\"|'[^']*')*?)(/?)>", re.S)
_ATTR = re.compile(r"([A-Za-z_][\w.-]*)\s*=\s*(\"([^\"]*)\"|'([^']*)'|([^\s>/]+))")


def parse_cml(text):
    """The element tree of a CML document (root element returned)."""
    text = re.sub(r"<!--.*?-->", "", text, flags=re.S)
    text = re.sub(r"<\?.*?\?>", "", text, flags=re.S)
    root = Element("#document", {})
    stack = [root]
    pos = 0
    for m in _TAG.finditer(text):
        stack[-1].text += text[pos:m.start()]
        pos = m.end()
        closing, name, attr_text, selfclose = m.group(1), m.group(2), m.group(3), m.group(4)
        if closing:
            while len(stack) > 1:
                top = stack.pop()
                if top.name == name:
                    break
            continue
        attrs = {a.group(1): next(g for g in a.groups()[2:] if g is not None) for a in _ATTR.finditer(attr_text)}
        el = Element(name, attrs)
        stack[-1].children.append(el)
        if not selfclose:
            stack.append(el)
    stack[-1].text += text[pos:]
    if not root.children:
        raise ValueError("no elements in the CML document")
    return root.children[0]


def _f(el, name, default=None):
    if el is None or name not in el.attrs:
        return default
    return float(el.attrs[name])


def _i(el, name, default=None):
    v = _f(el, name, None)
    return default if v is None else int(v)


def _b(el, name, default=False):
    if el is None or name not in el.attrs:
        return default
    return el.attrs[name].strip().lower() in ("true", "1")


# ------------------------------------------------------------------ data files

def _header_and_body(path):
    """Header lines (before the first line starting with '=') and the remaining text."""
    with _open_text(path) as f:
        lines = f.read().splitlines()
    for k, line in enumerate(lines):
        if line.startswith("="):
            return lines[:k], "\n".join(lines[k + 1:])
    return lines, ""


def _tags(header):
    """`[tag] value` pairs of a header (decodeLine, terrain_heightfield.cpp:88-127)."""
    out = {}
    for line in header:
        m = re.match(r"\s*(\[[^\]]*\])\s*(.*)$", line)
        if m and m.group(2).strip():
            out[m.group(1)] = m.group(2).strip()
    return out


def read_cbf(path):
    """(nx, ny, z) of a .cbf bathymetry file; z is (ny, nx) FP32, row j = y index."""
    header, body = _header_and_body(path)
    t = _tags(header)
    nx, ny = int(t["[nx]"]), int(t["[ny]"])
    z = np.array(body.split()[:nx * ny], dtype=np.float64).astype(np.float32)
    if z.size != nx * ny:
        raise ValueError(f"{path}: {z.size} heights for a {nx} x {ny} grid")
    return nx, ny, z.reshape(ny, nx)


def sample_corners(z, nx, ny, width, length):
    """Corner heights (ny+5, nx+5) for UpdateTerrainHeightfield: corner (i, j) at
    x = (i - 2.5)/(nx - 1) W - W/2, y = (j - 2.5)/(ny - 1) L, sampled bilinearly from the
    bathymetry grid with edge clamping (getBathymetry), all in FP32."""
    f = np.float32
    W, L = f(width), f(length)
    bny, bnx = z.shape
    i = np.arange(nx + 5, dtype=np.float32)
    j = np.arange(ny + 5, dtype=np.float32)
    x = (i - f(2.5)) / f(nx - 1) * W - (W / f(2))
    y = (j - f(2.5)) / f(ny - 1) * L
    xs = (x + W / f(2)) / W * f(bnx - 1)
    ys = y / L * f(bny - 1)
    xs = np.clip(xs, f(0), f(bnx - 1)).astype(np.float32)
    ys = np.clip(ys, f(0), f(bny - 1)).astype(np.float32)
    xd, xu = np.floor(xs).astype(np.int64), np.ceil(xs).astype(np.int64)
    yd, yu = np.floor(ys).astype(np.int64), np.ceil(ys).astype(np.int64)
    xf = (xs - xd.astype(np.float32)).astype(np.float32)
    yf = (ys - yd.astype(np.float32)).astype(np.float32)

    def lerp(t, a, b):                         # Lerp(x, down, up) = down + x (up - down)
        return (a + t * (b - a)).astype(np.float32)

    zd = lerp(xf[None, :], z[yd][:, xd], z[yd][:, xu])
    zu = lerp(xf[None, :], z[yu][:, xd], z[yu][:, xu])
    return lerp(yf[:, None], zd, zu)


def read_irregular_table(path):
    """Rows (amplitude, period, theta, phase) of an irregular-wave file (at most 1000)."""
    header, body = _header_and_body(path)
    t = _tags([l for l in header if not l.startswith("#")])
    n = min(MAX_IRREGULAR_WAVES, int(t.get("[NumberOfWaves]", "-1")))
    vals = np.array(body.split(), dtype=np.float64).astype(np.float32)
    rows = vals[:(vals.size // 4) * 4].reshape(-1, 4)
    return rows[:max(n, 0)].copy()


def read_time_series(path):
    """Rows (t, eta, hu, hv) of a uniform-time-series file."""
    _, body = _header_and_body(path)
    vals = np.array(body.split(), dtype=np.float64).astype(np.float32)
    return vals[:(vals.size // 4) * 4].reshape(-1, 4).copy()


def read_hot_start(path, nx, ny):
    """Ghost-inclusive (w, hu, hv) planes from a hot-start file (i j w hu hv alpha rows)."""
    with _open_text(path) as f:
        vals = np.array(f.read().split(), dtype=np.float64).astype(np.float32)
    X, Y = nx + 4, ny + 4
    rows = vals[:(vals.size // 6) * 6].reshape(-1, 6)
    if rows.shape[0] < X * Y:
        raise ValueError(f"{path}: {rows.shape[0]} cells for a {X} x {Y} grid")
    rows = rows[:X * Y]
    return np.stack([rows[:, 2].reshape(Y, X), rows[:, 3].reshape(Y, X), rows[:, 4].reshape(Y, X)])


# ------------------------------------------------------------------ experiment

@dataclass
class LogConfig:
    """<logData> (main.cpp:826-885): ranges and gauges are ghost-inclusive cell indices."""
    do_log: bool = False
    log_step: int = 1
    path: str = ""
    time_axis: str = "time"
    ranges: list = field(default_factory=list)     # (name, (x1, y1), (x2, y2))
    gauges_file: str = "gauges"
    gauges: list = field(default_factory=list)     # (x, y)


@dataclass
class Experiment:
    name: str
    s: object
    boundaries: list
    bottom: np.ndarray
    state: np.ndarray
    solitons: list
    log: LogConfig
    adaptive: dict
    visual: dict = field(default_factory=lambda: {"dissipation_threshold": 0.25, "whitewater_decay": 0.1})


def _resolve(base, p):
    p = p.strip()
    return p if os.path.isabs(p) else os.path.join(base, p)


def _boundary(el, base):
    """<westBoundary ...> etc. (main.cpp:656-825)."""
    if el is None:
        return Boundary(SOLID, 2)
    kind = el.attrs.get("type", "Solid").strip()
    width = int(_f(el, "widthNum", 2.0))
    level = _f(el, "seaLevel", 0.0)
    if kind == "Sponge":
        return Boundary(SPONGE, width, level)
    if kind == "SineWave":
        sw = el.child("sineWave")
        # PI / 180.0 * theta with the float PI promoted to double (main.cpp:684)
        theta = float(np.float32(float(PI_F) / 180.0 * float(np.float32(_f(sw, "theta", 0.0)))))
        return Boundary(SINEWAVE, width, level, amplitude=_f(sw, "amplitude", 0.0), period=_f(sw, "period", 1.0),
                        theta=theta)
    if kind == "IrregularWaves":
        path = _resolve(base, el.child("filePath").text)
        return Boundary(IRREGULARWAVES, width, level, waves=read_irregular_table(path))
    if kind == "UniformTimeSeries":
        path = _resolve(base, el.child("filePath").text)
        return Boundary(UNIFORMTIMESERIES, width, level, uts=read_time_series(path))
    return Boundary(SOLID, width, level)


def load_experiment(cml_path, gravity=9.81):
    """Settings, boundaries, bottom planes and initial state of a CML experiment, as the
    reference builds them at start-up (readInputCML, UpdateTerrainHeightfield,
    createSimTextures)."""
    base = os.path.dirname(os.path.abspath(cml_path))
    with _open_text(cml_path) as f:
        root = parse_cml(f.read())
    name_el = root.child("name")
    model = root.child("model")
    params = model.child("parameters") if model else None
    fric = model.child("friction") if model else None
    dims = root.child("fieldDimensions")
    grid = root.child("gridSize")
    nx, ny = _i(grid, "nx"), _i(grid, "ny")
    W, L = _f(dims, "width"), _f(dims, "length")
    swl = _f(dims, "stillWaterElevation", 0.0)
    s = settings(nx, ny, W, L, _f(params, "timestep", 0.0001), still_water_elevation=swl,
                 is_boussinesq=0 if (model is not None and model.attrs.get("type", "BSNQ").strip() == "NLSW") else 1,
                 epsilon=_f(params, "epsilon", 5e-8), theta=_f(params, "Theta", 1.3),
                 friction=_f(fric, "coef", 0.0),
                 is_manning=1 if (fric is not None and fric.attrs.get("type", "").strip() == "Manning") else 0,
                 correction_steps_num=_i(params, "correctionStepsNum", 0),
                 time_scheme=PREDICTOR_ADAPTIVE if _b(params, "adaptive") else PREDICTOR, gravity=gravity)
    adaptive = {"ema_alpha": _f(params, "exponentialMovingAverageAlpha", 0.5), "min_timestep_ratio": 0.05}
    bnd = [_boundary(root.child(f"{side}Boundary"), base) for side in ("west", "east", "south", "north")]
    # bathymetry at the corners, then the cell planes
    # a missing file gives a flat bed at 0 on the mesh's own grid (readBathymetry)
    bpath = root.child("bathymetryFilePath")
    bfile = _resolve(base, bpath.text) if bpath is not None else ""
    if bfile and (os.path.exists(bfile) or os.path.exists(bfile + ".gz")):
        _, _, z = read_cbf(bfile)
    else:
        z = np.zeros((ny, nx), np.float32)
    corners = sample_corners(z, nx, ny, W, L)
    bottom = bottom_from_corners(nx, ny, corners)
    # solitary waves: (H, theta in degrees, xc, yc, optional crest length)
    solitons = []
    for el in root.all("solitaryWave"):
        theta = float(PI_F / np.float32(180.0) * np.float32(_f(el, "theta", 0.0)))
        solitons.append((_f(el, "H", 0.0), theta, _f(el, "length", 1e30), _f(el, "xc", 0.0), _f(el, "yc", 0.0)))
    hot = root.child("hotStartFilePath")
    if hot is not None and os.path.exists(_resolve(base, hot.text)):
        state = hot_start_state(s, bottom, read_hot_start(_resolve(base, hot.text), nx, ny), solitons)
    else:
        state = initial_state(s, bottom, solitons)
    # logging
    log = LogConfig()
    ld = root.child("logData")
    if ld is not None:
        log.do_log = _b(ld, "doLog")
        log.log_step = _i(ld, "logStep", 1)
        lp = ld.child("logPath")
        log.path = lp.text.strip() if lp is not None else ""
        for el in ld.children:
            if el.name == "timeAxis":
                log.time_axis = el.attrs.get("filename", "time")
            elif el.name == "range":
                bl, tr = el.child("bottomLeft"), el.child("topRight")
                log.ranges.append((el.attrs.get("filename", "range"), (_i(bl, "x"), _i(bl, "y")),
                                   (_i(tr, "x"), _i(tr, "y"))))
            elif el.name == "gauges":
                log.gauges_file = el.attrs.get("filename", "gauges")
                nums = [int(v) for v in re.findall(r"-?\d+", el.text)]
                log.gauges = [(nums[k], nums[k + 1]) for k in range(0, len(nums) - 1, 2)]
    # <graphics><surfaceShading><dissipationIntensity threshold decay> (main.cpp:965-977)
    visual = {"dissipation_threshold": 0.25, "whitewater_decay": 0.1}
    gr = root.child("graphics")
    sh = gr.child("surfaceShading") if gr is not None else None
    di = sh.child("dissipationIntensity") if sh is not None else None
    if di is not None:
        visual = {"dissipation_threshold": _f(di, "threshold", 0.25), "whitewater_decay": _f(di, "decay", 0.1)}
    name = name_el.text.strip() if name_el is not None else os.path.basename(cml_path)
    return Experiment(name, s, bnd, bottom, state, solitons, log, adaptive, visual)
