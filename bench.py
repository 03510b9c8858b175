"""Benchmark of the Celeris Boussinesq step on one MI355X.

Workload: 2048 x 2048 interior cells, dx = 1 m, 1:30 plane from 8 m depth to 2 m dry land
plus a seeded Gaussian random field (std 0.3 m, correlation 20 cells), west irregular
waves (100 JONSWAP components, Hs 0.1 m, Tp 3 s, gamma 3.3, +-30 deg), 20-cell east
sponge, solid north/south walls, FP32.  A step is one full Boussinesq time step (explicit
stencil, x- and y-line tridiagonal solves, boundaries) over the whole grid; inputs are
resident in HBM before the timed region.

Prints one JSON line: Gcell-steps/s (interior cells x steps / wall time), the dominant
kernel's achieved HBM bandwidth against MI355X's 8 TB/s (and the library's own streaming
probes on the same GPU), the CPU oracle timed on a bounded sample of the same workload, a
bounded sample of the serving stream (scenarios with sides 256-2048 and 500-20000 steps,
the whole backlog at once and Poisson arrivals, continuous batching with all resident
requests advancing and with shortest-remaining-work-first) with p50/p99 completion latency
beside the plain engine serving the same stream one request at a time, the batched
ensemble (64 members of 128-1024 per side) against the same members one at a time, and the
plain 256^2 beach over its 2000 steps with its CPU-oracle baseline.

--workload ensemble prints the batched-ensemble line instead (64 scenarios with sides
log-uniform in [128, 1024], compensated update, 50-step graph blocks) against the same
scenarios run one at a time with the plain engine.  --workload precision runs the
precision case (4096 x 1024 solitary wave on a beach, 100000 steps) in FP32, compensated
FP32 and FP64 and reports each one's error against the FP64 run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL_STEP = 96          # algorithmic HBM bytes per interior cell-step (plain FP32)
# per state precision: the compensated state adds 12 B read + 12 B written (the low words),
# FP64 doubles every plane
PRECISIONS = {"fp32": (0, 96, "f32"), "compensated": (1, 120, "f32 + f32 (compensated)"), "fp64": (2, 192, "f64")}
HBM_PEAK_GBS = 8000.0             # MI355X HBM3E peak
# measured HBM bytes per explicit_step launch at 2048^2 (rocprofv3 FETCH_SIZE / WRITE_SIZE,
# scaled by the copy calibration) and the same kernel's VALU issue utilisation
PMC_FILE = os.path.join(ROOT, "profiles", "explicit_step_2048x2048_pmc.json")


def measured_traffic(nx, ny, band_rows):
    """HBM bytes per launch and VALU utilisation of the explicit kernel variant with this band
    height, from its rocprofv3 counter runs (None where none was profiled)."""
    if (nx, ny) != (2048, 2048) or not os.path.exists(PMC_FILE):
        return None, None
    with open(PMC_FILE) as f:
        d = json.load(f)
    v = d.get("by_band_rows", {}).get(str(band_rows))
    if v is None:
        return None, None
    return v["traffic_bytes_per_launch"], v.get("valu_busy")


def cpu_baseline(sc, seconds_target=15.0, max_steps=40):
    """FP32 CPU oracle (OpenMP over rows) on the same 2048^2 workload for a bounded number
    of steps; returns Gcell-steps/s and the sample description."""
    from oracle.binding import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    t0 = time.perf_counter()
    o.step(2)                                     # warm-up (Euler bootstrap steps)
    t1 = time.perf_counter()
    per = max((t1 - t0) / 2, 1e-6)
    n = int(max(3, min(max_steps, seconds_target / per)))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    cells = sc.s.nx * sc.s.ny
    return {"value": cells * n / dt / 1e9, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle, same {sc.s.nx}x{sc.s.ny} workload, {n} steps after 2 warm-up steps, "
                      f"{dt:.1f} s, OpenMP {threads} threads"}


def cpu_baselines(seconds_target=10.0):
    """The CPU oracles (FP32, FP64) single-threaded and with OpenMP on the host's cores, on the
    plain 256^2 beach case and the 2048^2 headline case.  Each run times as many steps as fit
    about seconds_target after 2 warm-up steps (the 256^2 case at most its full 2000 steps)."""
    from oracle.binding import Oracle, get_threads, set_threads
    from celeris_amd import scenarios as S
    all_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rows = []
    for name, sc, cap in (("plain 256x256", S.plain_beach(), 2000), ("headline 2048x2048", S.headline(), 100)):
        for prec in ("f32", "f64"):
            for threads in (1, all_threads):
                set_threads(threads)
                o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, prec)
                t0 = time.perf_counter()
                o.step(2)
                per = max((time.perf_counter() - t0) / 2, 1e-6)
                n = int(max(3, min(cap, seconds_target / per)))
                t0 = time.perf_counter()
                o.step(n)
                dt = time.perf_counter() - t0
                cells = sc.s.nx * sc.s.ny
                rows.append({"case": name, "precision": prec, "threads": get_threads(), "steps": n,
                             "seconds": round(dt, 2), "gcell_steps_per_s": cells * n / dt / 1e9})
                print(f"bench: cpu {name} {prec} {threads} threads: {rows[-1]['gcell_steps_per_s']:.4f} "
                      f"Gcell-steps/s ({n} steps)", file=sys.stderr, flush=True)
                del o
    set_threads(all_threads)
    return rows


def serving_leg(n, rates, seed, block):
    """A bounded sample of the serving stream: n requests (generated before the clock
    starts), served in real time with continuous batching with the whole backlog present at
    t = 0 (saturated) and at each Poisson arrival rate, and the plain engine serving them one
    at a time (FIFO replay of each request's measured time alone, setup included on both
    sides).  The same seed gives the same requests at every rate; arrival times scale with
    1/rate.  The bound is the largest ratio any batching could reach on this mix: the time
    alone summed over the requests against their cell-steps at the best single-grid rate."""
    from celeris_amd import ShallowWaterEngine
    from celeris_amd import scenarios as S
    from celeris_amd import serving
    base = S.serving_requests(n, 1.0, seed=seed)
    warm = S.small(nx=32, ny=32)                     # the HIP runtime's first-use cost is nobody's setup
    ShallowWaterEngine(warm.s, warm.boundaries, warm.bottom, warm.state).close()
    alone_times = {r.rid: serving.alone_run_time(r) for r in base}
    big = max(base, key=lambda r: r.scenario.s.nx * r.scenario.s.ny)
    e = ShallowWaterEngine(big.scenario.s, big.scenario.boundaries, big.scenario.bottom, big.scenario.state)
    e.timestep(20)
    e.sync()
    t0 = time.perf_counter()
    e.timestep(200)
    e.sync()
    best_rate = big.scenario.s.nx * big.scenario.s.ny * 200 / (time.perf_counter() - t0)
    e.close()
    sides = {}
    for r in base:
        sides[r.scenario.s.nx] = sides.get(r.scenario.s.nx, 0) + 1
    bound = serving.throughput_bound(base, alone_times, best_rate)
    out = {"workload": f"{n} requests, square grids {dict(sorted(sides.items()))}, 500-20000 steps (log-uniform), "
                       f"sloped seabed + GRF, sine/irregular forcing, compensated update, {block}-step blocks",
           "best_single_grid_gcell_steps_per_s": best_rate / 1e9, "throughput_ratio_bound": bound, "rates": {}}
    for rate in ["saturated"] + list(rates):
        if rate == "saturated":
            reqs = [S.Request(r.rid, 0.0, r.scenario, r.steps) for r in base]
            key = "saturated"
        else:
            reqs = [S.Request(r.rid, r.arrival / rate, r.scenario, r.steps) for r in base]
            key = f"{rate:g}/s"
        _, batched, _ = serving.serve(reqs, block=block)
        _, srpt, _ = serving.serve(reqs, block=block, policy="srpt")
        _, alone = serving.one_at_a_time(reqs, run_times=alone_times)
        out["rates"][key] = {
            "batched": batched, "batched_srpt": srpt, "one_at_a_time": alone,
            "throughput_ratio": batched["gcell_steps_per_s"] / alone["gcell_steps_per_s"],
            "throughput_ratio_srpt": srpt["gcell_steps_per_s"] / alone["gcell_steps_per_s"]}
        print(f"bench: serving {key}: batched/alone {out['rates'][key]['throughput_ratio']:.2f}, srpt "
              f"{out['rates'][key]['throughput_ratio_srpt']:.2f} (bound {bound:.2f})", file=sys.stderr, flush=True)
    return out


def device_bandwidth(nbytes=1 << 30, launches=20):
    """The library's HBM streaming probes (read, write, copy) on this GPU: GB/s of bytes moved."""
    from celeris_amd import probe_bandwidth
    return {k: round(probe_bandwidth(k, nbytes, launches)[1], 1) for k in ("read", "write", "copy")}


def stencil_pattern(nx, ny, band_rows, launches=20):
    """The explicit stencil's memory-access pattern without its arithmetic (the library's probe)
    in the stencil's layout and two others: microseconds per launch."""
    from celeris_amd import PATTERN_LAYOUTS, probe_stencil_pattern
    return {f"{k}_us": round(1000 * probe_stencil_pattern(k, nx, ny, band_rows, launches), 2) for k in PATTERN_LAYOUTS}


def solve_pattern(lines, launches=20):
    """The line solves' structure without their coefficient arithmetic (the library's probe):
    microseconds per launch over `lines` x- and y-lines of 2048 unknowns, equations combined by
    additions only and by cyclic reduction's eliminations."""
    from celeris_amd import probe_solve_pattern
    return {"additions_us": round(1000 * probe_solve_pattern(lines, False, launches), 2),
            "eliminations_us": round(1000 * probe_solve_pattern(lines, True, launches), 2)}


def bandwidth_line():
    from celeris_amd import probe_bandwidth
    rows = {}
    for nbytes in (256 << 20, 1 << 30, 4 << 30):
        for kind in ("read", "write", "copy"):
            ms, gbs = probe_bandwidth(kind, nbytes, 20)
            rows[f"{kind} {nbytes >> 20} MiB"] = {"ms": ms, "gb_per_s": gbs}
            print(f"bench: probe {kind} {nbytes >> 20} MiB: {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    best = max(v["gb_per_s"] for k, v in rows.items() if not k.endswith(" 256 MiB"))
    return {"metric": "HBM streaming rate (library probes)", "value": best, "unit": "GB/s", "n_gpus": 1,
            "steps": 20, "warmup": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "read / write / copy of 256 MiB, 1 GiB and 4 GiB device "
                                                     "buffers, 16 KiB per workgroup", "runs": rows,
                                          "peak_gb_per_s": HBM_PEAK_GBS}}


def ensemble_leg(n=64, steps=200, warmup=20, graph_steps=50):
    """The batched ensemble (n members, sides log-uniform in [128, 1024], compensated update) in
    captured blocks against the same members run one at a time with the plain engine: both
    sides time the same steps after the same warm-up, setup excluded from both."""
    from celeris_amd import Batch, ShallowWaterEngine
    from celeris_amd import scenarios as S
    members = S.ensemble(n, seed=4, steps=steps)
    cells = sum(m.s.nx * m.s.ny for m in members)
    b = Batch(graph_steps=graph_steps)
    for m in members:
        b.add(m.s, m.boundaries, m.bottom, m.state)
    b.step(warmup)
    rows = b.autotune(5, tune_lines=True)
    b.step(graph_steps)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    batched = cells * steps / (time.perf_counter() - t0)
    ok = b.failed() == []
    b.close()
    alone_s = 0.0
    for m in members:
        e = ShallowWaterEngine(m.s, m.boundaries, m.bottom, m.state)
        e.timestep(warmup)
        e.sync()
        t0 = time.perf_counter()
        e.timestep(steps)
        e.sync()
        alone_s += time.perf_counter() - t0
        e.close()
    alone = cells * steps / alone_s
    return {"workload": f"{n} scenarios, sides log-uniform in [128, 1024] ({cells} cells), compensated update, "
                        f"{graph_steps}-step captured blocks, {steps} timed steps",
            "batched_gcell_steps_per_s": batched / 1e9, "one_at_a_time_gcell_steps_per_s": alone / 1e9,
            "speedup": batched / alone, "band_rows": rows, "all_finite": ok}


def plain_leg(steps=2000):
    """The plain 256^2 beach for its 2000 steps on the GPU, and its CPU baseline: the FP32 oracle
    (OpenMP on the host's cores) over the same steps, which also checks the GPU's result."""
    from celeris_amd import ShallowWaterEngine
    from celeris_amd import scenarios as S
    from oracle.binding import Oracle, get_threads
    sc = S.plain_beach(steps=steps)
    cells = sc.s.nx * sc.s.ny
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(20)
    e.sync()
    ms = e.timed_steps(steps)
    e.close()
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(steps)
    a = e.get_state()
    e.close()
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f32")
    t0 = time.perf_counter()
    o.step(steps)
    cpu_s = time.perf_counter() - t0
    b = o.state().astype(np.float64)
    a = a.astype(np.float64)
    # max |difference| per field over the oracle's scale (eta: max |eta|; hu, hv: the larger
    # momentum component), and the wet cells (h > 1e-4 m) of both
    mom = max(np.abs(b[1]).max(), np.abs(b[2]).max())
    err = {"eta": float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max()),
           "hu": float(np.abs(a[1] - b[1]).max() / mom), "hv": float(np.abs(a[2] - b[2]).max() / mom)}
    wet = lambda st: (st[0] - sc.bottom[2]) > 1e-4
    return {"workload": f"plain beach {sc.s.nx}x{sc.s.ny}, dx 0.5 m, west sine, walls, east sponge, {steps} steps",
            "gpu_gcell_steps_per_s": cells * steps / (ms * 1e-3) / 1e9,
            "cpu_oracle_gcell_steps_per_s": cells * steps / cpu_s / 1e9, "cpu_threads": get_threads(),
            "relative_error_vs_oracle": err,
            "same_wet_dry_cells": bool((wet(a) == wet(b)).all())}


def ensemble_line(args):
    from celeris_amd import Batch, ShallowWaterEngine
    from celeris_amd import scenarios as S
    members = S.ensemble(args.members, seed=4, steps=args.steps)
    cells = sum(m.s.nx * m.s.ny for m in members)
    b = Batch(graph_steps=args.graph_steps)
    ids = [b.add(m.s, m.boundaries, m.bottom, m.state) for m in members]
    b.step(args.warmup)
    band_rows = b.autotune(args.autotune, tune_lines=True) if args.autotune > 0 else None
    b.step(10)
    b.sync()
    t0 = time.perf_counter()
    ms = b.timed_steps(args.steps)
    wall = time.perf_counter() - t0
    finite = all(bool(np.isfinite(b.get_state(i)).all()) for i in ids)
    b.close()
    batched = cells * args.steps / wall / 1e9
    # the same members one at a time with the plain engine (step-by-step launches)
    alone_wall = 0.0
    for m in members:
        e = ShallowWaterEngine(m.s, m.boundaries, m.bottom, m.state)
        e.timestep(args.warmup)
        e.sync()
        t0 = time.perf_counter()
        e.timed_steps(args.steps)
        alone_wall += time.perf_counter() - t0
        e.close()
    alone = cells * args.steps / alone_wall / 1e9
    sides = [(m.s.nx, m.s.ny) for m in members]
    return {
        "metric": "Gcell-steps/s (batched ensemble)", "value": batched, "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(members)} scenarios, sides log-uniform in [128, 1024] "
                               f"({sum(nx * ny for nx, ny in sides)} cells), dx 0.5 m, sloped seabed + GRF, "
                               "sine/irregular forcing, compensated update",
                   "graph_steps": args.graph_steps, "event_ms_per_step": ms / args.steps,
                   "autotuned_band_rows": band_rows,
                   "one_at_a_time_gcell_steps_per_s": alone, "speedup_vs_one_at_a_time": batched / alone,
                   "state_finite": finite, "smallest": min(sides, key=lambda v: v[0] * v[1]),
                   "largest": max(sides, key=lambda v: v[0] * v[1])},
    }


def precision_variants(nx, ny, steps):
    """The precision case: nx x ny channel (dx 0.25 m, 1:50 beach to +0.5 m, solitary wave
    H/h = 0.3, walls), run in FP32, compensated FP32 and FP64 on the GPU.  The FP64 engine is
    the reference (it equals the FP64 CPU oracle to round-off at sizes the oracle can run);
    each variant reports its eta error over H at the end, its maximum run-up on the centre
    row, its volume drift and its throughput."""
    from celeris_amd import STATE_COMPENSATED, STATE_FP32, STATE_FP64, ShallowWaterEngine
    from celeris_amd import scenarios as S
    H, chunk = 0.3, 2000
    sc = S.solitary_channel(nx=nx, ny=ny, dx=0.25, steps=steps)
    row = sc.s.ny // 2 + 2
    cells = sc.s.nx * sc.s.ny
    xy = [(i, row) for i in range(sc.s.nx + 4)]
    ba = sc.bottom[2][row].astype(np.float64)
    out, finals = {}, {}
    for name, prec in (("fp64", STATE_FP64), ("fp32", STATE_FP32), ("compensated", STATE_COMPENSATED)):
        sc.s.state_precision = prec
        e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
        v0 = e.stats()["sum_h"]
        runup, wall, done = -1e9, 0.0, 0
        while done < steps:
            k = min(chunk, steps - done)
            t0 = time.perf_counter()
            e.timestep(k)
            e.sync()
            wall += time.perf_counter() - t0
            done += k
            w = e.sample(xy)[:, 0].astype(np.float64)
            wet = (w - ba) > 1e-3
            if wet.any():
                runup = max(runup, float(ba[wet].max()))
            print(f"bench: precision {name} {done}/{steps}", file=sys.stderr, flush=True)
        finals[name] = e.get_state_f64()
        out[name] = {"gcell_steps_per_s": cells * steps / wall / 1e9, "runup_m": runup,
                     "volume_drift": (e.stats()["sum_h"] - v0) / v0}
        e.close()
    ref = finals["fp64"][0][2:-2, 2:-2]
    for name in ("fp32", "compensated", "fp64"):
        d = finals[name][0][2:-2, 2:-2] - ref
        out[name]["eta_linf_over_H"] = float(np.abs(d).max() / H)
        out[name]["eta_l2_over_H"] = float(np.sqrt((d ** 2).mean()) / H)
    workload = (f"solitary wave H/h 0.3 on a 1:50 beach, {sc.s.nx}x{sc.s.ny} cells, dx 0.25 m, "
                f"dt {sc.s.timestep:.5f} s, walls, {steps} steps")
    return workload, out


def precision_line(args):
    """The precision case at its full size and length (4096 x 1024, 100 000 steps)."""
    workload, out = precision_variants(args.nx, args.ny, args.steps)
    return {"metric": "precision case (eta error vs GPU FP64)", "value": out["compensated"]["eta_linf_over_H"],
            "unit": "L-inf eta error / H (compensated)", "n_gpus": 1, "steps": args.steps, "warmup": 0,
            "higher_is_better": False, "vs_baseline": None, "dtype": "f32 / f32+f32 / f64", "data": "synthetic",
            "config": {"workload": workload, "variants": out}}


def precision_leg(steps=10000):
    """The precision case at its full size for a bounded number of steps: the three GPU
    variants against the GPU FP64 engine."""
    workload, out = precision_variants(4096, 1024, steps)
    return {"workload": workload, "variants": out}


def precision_oracle_check(oracle_steps=20):
    """CPU baseline of the precision case: the GPU FP64 engine against the CPU FP64 oracle
    (OpenMP) over oracle_steps steps of the 4096 x 1024 channel, and the oracle's rate."""
    from celeris_amd import STATE_FP64, ShallowWaterEngine
    from celeris_amd import scenarios as S
    from oracle.binding import Oracle, get_threads
    sc = S.solitary_channel(nx=4096, ny=1024, dx=0.25, steps=oracle_steps)
    sc.s.state_precision = STATE_FP64
    e = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    e.timestep(oracle_steps)
    a = e.get_state_f64()
    e.close()
    o = Oracle(sc.s, sc.boundaries, sc.bottom, sc.state, "f64")
    t0 = time.perf_counter()
    o.step(oracle_steps)
    cpu_s = time.perf_counter() - t0
    b = o.state().astype(np.float64)
    scale = [np.abs(b[0]).max(), max(np.abs(b[1]).max(), np.abs(b[2]).max())]
    err = {"eta": float(np.abs(a[0] - b[0]).max() / scale[0]), "hu": float(np.abs(a[1] - b[1]).max() / scale[1]),
           "hv": float(np.abs(a[2] - b[2]).max() / scale[1])}
    return {"steps": oracle_steps, "relative_error": err,
            "cpu_oracle_gcell_steps_per_s": 4096 * 1024 * oracle_steps / cpu_s / 1e9, "cpu_threads": get_threads()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="steps per captured hipGraph block (0: launch step by step, the faster mode measured on "
                         "MI355X; the ensemble workload defaults to 50-step blocks)")
    ap.add_argument("--autotune", type=int, default=30,
                    help="repeats per candidate when autotuning band height and line-solve shapes (0: defaults); "
                         "candidates within a few per cent of each other need about 30 to be told apart")
    ap.add_argument("--band-rows", type=int, default=0, choices=[0, 2, 4, 8, 12, 16],
                    help="pin the explicit kernel's band height after autotuning (0: keep the autotuned one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-steps", type=int, default=10000,
                    help="steps of the 4096x1024 precision case in the default line (0: skip)")
    ap.add_argument("--precision", choices=list(PRECISIONS), default="fp32",
                    help="state precision of the headline workload (the headline figure is fp32)")
    ap.add_argument("--workload", choices=["headline", "ensemble", "precision", "cpu", "bandwidth", "serving"], default="headline")
    ap.add_argument("--members", type=int, default=64, help="ensemble scenarios")
    ap.add_argument("--serving-requests", type=int, default=24,
                    help="requests in the serving sample of the headline line (0: skip)")
    ap.add_argument("--serving-rates", type=str, default="5,50", help="Poisson arrival rates (1/s), comma-separated")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the batched-ensemble leg of the headline line (0: skip)")
    ap.add_argument("--probe", choices=["read", "write", "copy"], default=None,
                    help="only run the library's HBM streaming probe of this kind on a 1 GiB buffer (4 launches: the "
                         "known-size reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.probe:
        from celeris_amd import probe_bandwidth
        ms, gbs = probe_bandwidth(args.probe, 1 << 30, 4)
        print(json.dumps({"probe": args.probe, "bytes_per_launch": 1 << 30, "launches": 5, "ms_per_launch": ms,
                          "gb_per_s": gbs}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of the step is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "precision":
        if "--steps" not in sys.argv:
            args.steps = 100000
        if "--nx" not in sys.argv:
            args.nx, args.ny = 4096, 1024
        print(json.dumps(precision_line(args)), flush=True)
        return
    if args.workload == "serving":
        rates = [float(v) for v in args.serving_rates.split(",")]
        leg = serving_leg(args.serving_requests, rates, 5, args.graph_steps or 50)
        sat = leg["rates"]["saturated"]
        print(json.dumps({"metric": "serving throughput ratio (batched / one at a time, saturated)",
                          "value": sat["throughput_ratio"], "unit": "x", "n_gpus": 1, "steps": 0, "warmup": 0,
                          "higher_is_better": True, "vs_baseline": None, "dtype": "f32 + f32 (compensated)",
                          "data": "synthetic", "config": {"workload": leg["workload"], "serving": leg}}), flush=True)
        return
    if args.workload == "bandwidth":
        print(json.dumps(bandwidth_line()), flush=True)
        return
    if args.workload == "cpu":
        rows = cpu_baselines()
        head = [r for r in rows if r["case"].startswith("headline") and r["precision"] == "f32"][-1]
        print(json.dumps({"metric": "CPU oracle baselines (reported only)", "value": head["gcell_steps_per_s"],
                          "unit": "Gcell-steps/s", "n_gpus": 0, "steps": head["steps"], "warmup": 2,
                          "higher_is_better": True, "vs_baseline": None, "dtype": "f32 / f64", "data": "synthetic",
                          "config": {"workload": "CPU FP32 and FP64 oracles, 1 thread and OpenMP, plain 256x256 "
                                                 "beach and 2048x2048 headline", "runs": rows}}), flush=True)
        return
    if args.workload == "ensemble":
        if args.graph_steps == 0 and "--graph-steps" not in sys.argv:
            args.graph_steps = 50
        print(json.dumps(ensemble_line(args)), flush=True)
        return

    from celeris_amd import ShallowWaterEngine
    from celeris_amd import scenarios as S

    sc = S.headline(nx=args.nx, ny=args.ny)
    prec_code, bytes_per_cell, dtype = PRECISIONS[args.precision]
    sc.s.state_precision = prec_code
    eng = ShallowWaterEngine(sc.s, sc.boundaries, sc.bottom, sc.state)
    cells = sc.s.nx * sc.s.ny
    eng.set_graph_steps(args.graph_steps)
    eng.timestep(args.warmup)
    tuning = eng.autotune(args.autotune) if args.autotune > 0 else eng.tuning()
    if args.band_rows:
        eng.set_tuning(args.band_rows, tuning["x_chunk"], tuning["x_waves"], tuning["y_chunk"], tuning["y_waves"])
        tuning = eng.tuning()
    eng.timestep(10)
    eng.sync()

    # timed region: K steps bracketed by device synchronisation, wall clock and HIP events
    t0 = time.perf_counter()
    ms_events = eng.timed_steps(args.steps)
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall / 1e9

    # the same number of steps again with per-kernel HIP events on the engine's stream
    eng.set_profiling(True)
    eng.timed_steps(args.steps)
    kern = {}
    for name in ("explicit_step", "solve_lines", "boundary_apply"):
        ms, n = eng.kernel_time(name)
        kern[name] = ms / max(n, 1)
    eng.set_profiling(False)
    st = eng.get_state()
    finite = bool(np.isfinite(st).all())
    eng.close()

    avg_ms = kern["explicit_step"]
    achieved = bytes_per_cell * cells / (avg_ms * 1e-3) / 1e9
    traffic, valu_busy = (measured_traffic(sc.s.nx, sc.s.ny, tuning["band_rows"]) if args.precision == "fp32"
                          else (None, None))
    out = {
        "metric": "Gcell-steps/s (2048^2 Boussinesq)",
        "value": value,
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"Boussinesq step, {sc.s.nx}x{sc.s.ny} interior cells, dx 1 m, 1:30 slope + GRF, "
                               "west irregular waves (100 JONSWAP components), 20-cell east sponge, N/S walls, "
                               + args.precision.upper() + " state",
                   "nx": sc.s.nx, "ny": sc.s.ny, "dt": sc.s.timestep, "graph_steps": args.graph_steps,
                   "tuning": {k: (round(v, 2) if isinstance(v, float) else v) for k, v in tuning.items()},
                   "hbm_frac_step": value * 1e9 * bytes_per_cell / (HBM_PEAK_GBS * 1e9),
                   "event_ms_per_step": ms_events / args.steps,
                   "kernel_avg_ms": kern, "state_finite": finite},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"explicit_step (AB3, {tuning['band_rows']}-row bands"
                               + (", fast variant)" if tuning.get("fast_arith") else ")"),
                     "bytes_per_launch": bytes_per_cell * cells,
                     "measured_hbm_gbs": None if traffic is None else traffic / (avg_ms * 1e-3) / 1e9,
                     "valu_busy": valu_busy},
        "cpu_baseline": None,
    }
    print(f"bench: {sc.s.nx}x{sc.s.ny} {args.precision} step {value:.2f} Gcell-steps/s", file=sys.stderr, flush=True)
    out["device_bandwidth_gb_per_s"] = device_bandwidth()
    print(f"bench: probes {out['device_bandwidth_gb_per_s']}", file=sys.stderr, flush=True)
    if args.precision == "fp32" and tuning["band_rows"] in (8, 12, 16):
        # the stencil's data movement alone, same grid and band height, beside its measured time
        pat = stencil_pattern(sc.s.nx, sc.s.ny, tuning["band_rows"])
        pat["explicit_step_us"] = round(1000 * avg_ms, 2)
        out["roofline"]["access_pattern_probe"] = pat
        print(f"bench: stencil access pattern alone {pat}", file=sys.stderr, flush=True)
    if args.precision == "fp32" and sc.s.nx == 2048 and sc.s.ny == 2048:
        # the line solves' loads, exchanges, barriers and stores alone, beside their measured time
        sp = solve_pattern(2048)
        sp["solve_lines_us"] = round(1000 * kern["solve_lines"], 2)
        out["line_solve_probe"] = sp
        print(f"bench: line-solve structure alone {sp}", file=sys.stderr, flush=True)
    if args.serving_requests > 0:
        rates = [float(v) for v in args.serving_rates.split(",")]
        out["serving"] = serving_leg(args.serving_requests, rates, 5, args.graph_steps or 50)
        print("bench: serving sample done", file=sys.stderr, flush=True)
    if args.ensemble_members > 0:
        out["ensemble"] = ensemble_leg(args.ensemble_members)
        print(f"bench: ensemble {out['ensemble']['batched_gcell_steps_per_s']:.2f} Gcell-steps/s batched, "
              f"{out['ensemble']['speedup']:.2f}x one at a time", file=sys.stderr, flush=True)
    if args.precision_steps > 0:
        out["precision"] = precision_leg(args.precision_steps)
        v = out["precision"]["variants"]
        print(f"bench: precision eta Linf/H fp32 {v['fp32']['eta_linf_over_H']:.2e}, compensated "
              f"{v['compensated']['eta_linf_over_H']:.2e}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        if args.precision_steps > 0:
            out["precision"]["fp64_vs_cpu_fp64_oracle"] = precision_oracle_check()
            print(f"bench: precision FP64 vs CPU FP64 oracle {out['precision']['fp64_vs_cpu_fp64_oracle']['relative_error']}",
                  file=sys.stderr, flush=True)
        out["plain_256"] = plain_leg()
        print(f"bench: plain 256^2 {out['plain_256']['gpu_gcell_steps_per_s']:.2f} Gcell-steps/s", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
