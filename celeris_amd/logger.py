"""Logging of a running engine in the reference's text formats (st_DumpToFile,
engine.cpp:628-675; dumpInundationToFile, engine.cpp:781-817; dumpBathymetryToFile,
engine.cpp:750-779):

  <time_axis>.txt          one line per record: the simulated time
  <time_axis>-elapsed.txt  one line per record: wall-clock seconds since the logger started
  <range name>.txt         per record, every cell of the range: i j w hu hv 0
  <gauges file>.txt        per record, every gauge: i j w hu hv 0
  <date>-inundation.txt    the maximum inundated depth of every interior cell (the
                           auxiliary texture's r channel, kept on the device by the engine's
                           visual outputs)
  <date>.cbf               the cell-average bottom BA of the interior cells

Records follow the reference's schedule (engine.cpp:1488-1495): after a step that ends a
predictor cycle (the next scheme is Euler or a predictor), when the step counter
timestep_count (before its increment) is a multiple of logStep; the time written is the
virtual time before that step's dt is added.  Ranges and gauges use ghost-inclusive cell
indices; cells outside the (nx+4) x (ny+4) grid are left out with a warning (the
reference reads past its texture there).  Gauges are read with cel_engine_sample;
ranges download the state.
"""
import os
import time
import warnings

from . import CORRECTOR


def _date_stamp():
    t = time.localtime()       # year-month-day hour-minute-second, unpadded (engine.cpp:757-758)
    return f"{t.tm_year}-{t.tm_mon}-{t.tm_mday} {t.tm_hour}-{t.tm_min}-{t.tm_sec}"


class Logger:
    def __init__(self, engine, log_cfg, out_dir=None, inundation=True, dissipation_threshold=0.25,
                 whitewater_decay=0.1):
        self.e = engine
        self.cfg = log_cfg
        self.dir = out_dir or log_cfg.path or "."
        os.makedirs(self.dir, exist_ok=True)
        X, Y = engine.nx + 4, engine.ny + 4
        inside = lambda i, j: 0 <= i < X and 0 <= j < Y
        self.gauges = [g for g in log_cfg.gauges if inside(*g)]
        self.ranges = [r for r in log_cfg.ranges if inside(*r[1]) and inside(*r[2])]
        self.skipped = [g for g in log_cfg.gauges if not inside(*g)] + [r for r in log_cfg.ranges if r not in self.ranges]
        if self.skipped:
            warnings.warn(f"log cells outside the {X} x {Y} grid left out: {self.skipped}")
        self.inundation = inundation
        if inundation:
            engine.set_visual(True, dissipation_threshold, whitewater_decay)
        self.t0 = time.perf_counter()

    def _append(self, name, lines):
        with open(os.path.join(self.dir, name + ".txt"), "a") as f:
            f.write("".join(lines))

    def record(self, sim_time):
        """One log record of the current state, stamped with sim_time."""
        self._append(self.cfg.time_axis, [f"{sim_time:g}\n"])
        self._append(self.cfg.time_axis + "-elapsed", [f"{time.perf_counter() - self.t0:g}\n"])
        if self.ranges:
            st = self.e.get_state()
            for name, (x1, y1), (x2, y2) in self.ranges:
                lines = []
                for j in range(y1, y2 + 1):
                    for i in range(x1, x2 + 1):
                        lines.append(f"{i}\t{j}\t{st[0, j, i]:g}\t{st[1, j, i]:g}\t{st[2, j, i]:g}\t0\n")
                self._append(name, lines)
        if self.gauges:
            vals = self.e.sample(self.gauges)
            lines = [f"{i}\t{j}\t{v[0]:g}\t{v[1]:g}\t{v[2]:g}\t0\n" for (i, j), v in zip(self.gauges, vals)]
            self._append(self.cfg.gauges_file, lines)

    def run(self, steps):
        """Steps the engine `steps` times, writing a record wherever the reference would."""
        log_step = self.cfg.log_step if self.cfg.do_log else 0
        for _ in range(steps):
            t_before = self.e.time
            count = self.e.timestep_count
            self.e.timestep(1)
            if log_step and self.e.scheme != CORRECTOR and count % log_step == 0:
                self.record(t_before)

    def dump_inundation(self):
        """Writes the maximum inundated depth of every interior cell; returns the file name."""
        if not self.inundation:
            raise RuntimeError("the logger was created without the inundation record")
        depth = self.e.aux()[0]
        nx, ny = self.e.nx, self.e.ny
        name = _date_stamp() + "-inundation.txt"
        with open(os.path.join(self.dir, name), "w") as f:
            f.write(f"[nx] {nx}\n[ny] {ny}\n\n\n")
            f.write("# This file contains the maximum depth of inundation on each point, in meters.\n\n")
            f.write("====================================\n")
            for j in range(2, ny + 2):
                f.write("".join(f"{v:g}\t" for v in depth[j, 2:nx + 3]) + "\n")
        return name

    def dump_bathymetry(self, bottom):
        """Writes the interior cell-average bottom (BA) as a .cbf file; returns the file name."""
        nx, ny = self.e.nx, self.e.ny
        name = _date_stamp() + ".cbf"
        with open(os.path.join(self.dir, name), "w") as f:
            f.write(f"[nx] {nx}\n[ny] {ny}\n\n\n====================================\n")
            for j in range(2, ny + 2):
                f.write("".join(f"{v:g}\t" for v in bottom[2][j, 2:nx + 2]) + "\n")
        return name
