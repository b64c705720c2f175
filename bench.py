#!/usr/bin/env python3
"""Benchmark of the SW2D time step on one MI355X.

Workload of the headline line: the 2048 x 2048 "plain FP64 GPU drop-in at
real size" case -- dx = 5 m, 0.5 % slope + 3 Gaussian-valley channels + fBm
3 m + prefilled Gaussian lake pit whose rim cells are the EFDC inflow records,
Manning 0.02 (channels) / 0.08 (hillslopes), moving Gaussian storm (sigma
1.5 km, 80 mm/h, 5 m/s, 1 h), CFL-adaptive step (Courant 0.5) -- run through
the fused FP64 kernel.  One step = one loop iteration of the reference (flux,
continuity, treat_error, continuity, derived fields) over all cells; the
metric is cell-updates per second with the state resident in HBM.

Also measured (same JSON line): the FP32-state / FP64-balance form on the same
case; the plain FP64 kernel sequence; a batch of 256 ensemble scenarios
(128..1024 square grids, 1 h each) stepped together; a bounded serving
session (Poisson arrivals, grids 256..4096, 2 h storms) with p50/p99
time-to-hydrograph; the plain FP64 path's time for the same scenarios one at
a time; the CPU baseline (the FP64 oracle, OpenMP, on a bounded sample).
"""
import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
METRIC = "cell-updates/s + % HBM roofline; p50/p99 scenario latency vs plain FP64 GPU"
TRAFFIC_TABLE = ROOT / "profiles" / "fused_step_2048_dropin_traffic.json"
BATCH_TRAFFIC = ROOT / "profiles" / "fused_step_batch256_traffic.json"
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "sw2d_efdc_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = ROOT / "sw2d-efdc_amd"
    spec = importlib.util.spec_from_file_location(name, d / "__init__.py",
                                                  submodule_search_locations=[str(d)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    """Progress on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def bytes_per_cell(precision, has_class_map):
    """Compulsory HBM bytes of one cell-update of the fused kernel: read h, um,
    vn and the FP64 bed z, write h, um, vn (+1 byte Manning class when a class
    map is used): 57 B in FP64, 33 B in the FP32 form."""
    e = 8 if precision == "fp64" else 4
    return 6 * e + 8 + (1 if has_class_map else 0)


def make_case(pkg, spec, precision, kernel, tile_rows):
    c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                 precision=precision, kernel=kernel, adaptive=spec.adaptive,
                 courant=spec.courant, tile_rows=tile_rows)
    if spec.rain is not None:
        c.set_rain(spec.rain, spec.dtrain)
    if spec.nclass is not None:
        c.set_manning_map(spec.nclass, spec.ntable)
    if spec.storm is not None:
        s = spec.storm
        c.set_storm(s["peak"], s["x0"], s["y0"], s["vx"], s["vy"], s["sigma"], s["t0"], s["t1"])
    if spec.outlet is not None:
        c.set_outlet(*spec.outlet)
    if spec.lake_mask is not None:
        c.set_lake_bc(spec.lake_mask)
    return c


def measured_traffic(precision, tile_rows, grid):
    """HBM bytes per launch of the fused kernel at this tile height, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE table committed under profiles/ (None
    when the table does not cover this case)."""
    try:
        t = json.load(open(TRAFFIC_TABLE))
    except (OSError, ValueError):
        return None
    if list(t.get("grid", [])) != list(grid):
        return None
    return t.get("bytes_per_launch", {}).get(precision, {}).get(str(tile_rows))


def measured_traffic_mix(precision, heights, grid):
    """Launch-weighted mean of measured_traffic over (iterations, tile rows)
    chunks (None unless the table covers every height that ran)."""
    vals = [(n, measured_traffic(precision, H, grid)) for n, H in heights]
    if not vals or any(t is None for _, t in vals):
        return None
    return int(round(sum(n * t for n, t in vals) / sum(n for n, _ in vals)))


def measured_batch_traffic(tile_rows, cells):
    """HBM bytes per launch of the batch leg's launch at this tile height, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE table of the same batch
    (None for another ensemble or a height the table lacks)."""
    try:
        t = json.load(open(BATCH_TRAFFIC))
    except (OSError, ValueError):
        return None
    if t.get("cells") != cells:
        return None
    per = t.get("bytes_per_launch")
    if isinstance(per, dict):
        return per.get(str(tile_rows))
    return per if t.get("tile_rows") == tile_rows else None


def timed_run(pkg, spec, precision, kernel, tile_rows, warmup, steps, retune=0):
    """Autotune (tile_rows 0), warm up, then time `steps` iterations with HIP
    events on the library's stream around every launch (or graph replay).
    With retune > 0 the library re-times the tile height every `retune`
    iterations inside the timed steps (its cost is in the wall time)."""
    case = make_case(pkg, spec, precision, kernel, tile_rows)
    case.step(warmup)
    if kernel == "fused" and tile_rows == 0:
        tile_rows = case.autotune(10)  # on the warmed-up state
        if retune > 0:
            case.set_retune(retune, 3)
    case.sync()
    case.timing(True)
    case.sync()
    # stepped in retune-interval chunks (no extra synchronisation): a retune
    # happens at the start of a chunk, so each chunk runs at one tile height
    heights = []
    t0 = time.perf_counter()
    left = steps
    while left > 0:
        n = min(left, retune) if (retune > 0 and kernel == "fused") else left
        case.step(n)
        heights.append((n, case.tile_rows))
        left -= n
    case.sync()
    wall = time.perf_counter() - t0
    kms, launches = case.kernel_time()
    case.timing(False)
    R, C = spec.z.shape
    t_sim, _, dt2 = case.clock()
    bpc = bytes_per_cell(precision, spec.nclass is not None)
    avg_s = (kms / 1e3) / max(launches, 1)
    return case, dict(value=R * C * steps / wall, ms_per_step=1e3 * wall / steps,
                      tile_rows=tile_rows, heights=heights, t_sim=t_sim, dt2=dt2, balance=case.balance(), bpc=bpc,
                      avg_s=avg_s, achieved=bpc * R * C / avg_s / 1e9,
                      ws_mb=case.device_bytes() / 1e6, device_bytes=case.device_bytes())


def plain_rate(pkg, side, rates, iters=30):
    """cell-updates/s of the plain FP64 kernel sequence on one side x side
    ensemble scenario (cached per side)."""
    if side in rates:
        return rates[side]
    cases = sys.modules["sw2d_efdc_amd.cases"]
    spec = cases.ensemble_scenario(0, side=side)
    c = make_case(pkg, spec, "fp64", "plain", 0)
    c.step(3)
    c.sync()
    t0 = time.perf_counter()
    c.step(iters)
    c.sync()
    rates[side] = side * side * iters / (time.perf_counter() - t0)
    c.close()
    return rates[side]


def plain_time(pkg, cells_iters, rates):
    """Seconds the plain FP64 path takes for (side, iterations) scenarios run one
    at a time, from its measured rate at the nearest measured side."""
    total = 0.0
    for side, iters in cells_iters:
        key = min((128, 256, 512, 1024, 2048, 4096), key=lambda s: abs(np.log(s / side)))
        total += side * side * iters / plain_rate(pkg, key, rates)
    return total


def batch_leg(pkg, cases, n, plain_rates):
    """The ensemble: n scenarios (128..1024, 1 h, fixed dt 0.5 s) in one batch,
    FP32 state / FP64 balance, tile rows by sw2d_batch_autotune after 300 iterations."""
    specs = [cases.ensemble_scenario(k) for k in range(n)]
    log(f"batch: {n} scenarios generated")
    need = sum(pkg.footprint(s.z.shape[0], s.z.shape[1], precision="fp32", batch_tile_rows=8,
                             hydro_cap=int(s.t_end / (2 * s.dt)) + 16) for s in specs)
    b = pkg.Batch(precision="fp32", pool_bytes=int(need * 1.1) + (64 << 20))
    members = []
    t_setup = time.perf_counter()
    for s in specs:
        m = b.case(s.z, s.h0, s.cellsize, dt=s.dt, manning=s.manning)
        m.set_rain(s.rain, s.dtrain)
        m.set_outlet(*s.outlet)
        m.set_hydrograph(int(s.t_end / (2 * s.dt)) + 16)
        b.attach(m, t_end=s.t_end)
        members.append(m)
    t_setup = time.perf_counter() - t_setup
    b.timing(True)
    t0 = time.perf_counter()
    # the tile height is tuned once the rain has wetted the slopes (on the dry
    # initial state the flux work is skipped and the timing unrepresentative);
    # the trial launches are inside the timed region
    b.step(300)
    H = b.autotune(8)
    while b.step(60):
        pass
    wall = time.perf_counter() - t0
    kms, launches = b.kernel_time()
    iters = [b.status(m)[1] for m in members]
    cell_updates = sum(s.cells * k for s, k in zip(specs, iters))
    bytes_alg = sum(s.cells * k for s, k in zip(specs, iters)) * bytes_per_cell("fp32", False)
    achieved = bytes_alg / (kms / 1e3) / 1e9
    vol_err = max(abs(m.balance()["error"]) / max(m.balance()["volume"], 1e-30) for m in members)
    q_peak = [m.balance()["q_peak"] for m in members]
    t_plain = plain_time(pkg, [(s.z.shape[0], k) for s, k in zip(specs, iters)], plain_rates)
    b.close()
    return {
        "workload": f"{n} ensemble scenarios, square grids 128..1024 (log-uniform), dx 10 m, "
                    "slope U(0.2%,3%), fBm U(0.5,5) m, n U(0.02,0.1), block rain U(10,120) mm/h "
                    "for U(15,90) min, free outlet, fixed dt 0.5 s, 1 h each, one launch per "
                    "iteration for all",
        "precision": "f32 state / f64 balance",
        "cells_total": int(sum(s.cells for s in specs)),
        "tile_rows": H,
        "iterations": int(max(iters)),
        "wall_s": wall,
        "setup_s": t_setup,
        "value": cell_updates / wall,
        "unit": "cell-updates/s",
        "scenarios_per_s": n / wall,
        "roofline": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_batch_traffic(H, int(sum(s.cells for s in specs))),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE of this batch; "
                                       "profiles/fused_step_batch256_ensemble.md",
                     "bytes_per_cell_update": bytes_per_cell("fp32", False),
                     "kernel_ms_total": kms, "launches": launches},
        "max_volume_error_rel": vol_err,
        "mean_q_peak_m3s": float(np.mean(q_peak)),
        "plain_fp64_one_at_a_time_s": t_plain,
        "scenario_throughput_vs_plain_fp64": t_plain / wall,
    }


def serve_leg(pkg, cases, seconds, rate, hours, plain_rates, active_cells=0, tile=0):
    """A bounded serving session: Poisson arrivals at `rate`/s for `seconds`,
    grid sides 256..4096 (p 0.4/0.3/0.15/0.1/0.05), 2 h storms, FP32 form."""
    n = max(1, int(round(seconds * rate)))
    arrivals, specs = cases.serve_requests(n, rate=rate, t_end=3600.0 * hours)
    log(f"serving: {n} requests generated; submitting")
    srv = pkg.Server(precision="fp32", pool_bytes=64 << 30, tile_rows=tile, slice_iters=60,
                     max_active_cells=active_cells)
    cap = int(3600.0 * hours / (2 * 0.5)) + 16
    ids = []
    t0 = time.perf_counter()
    for a, s in zip(arrivals, specs):
        d = a - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        ids.append(srv.submit(s.z, s.cellsize, t_end=s.t_end, rain=s.rain, dtrain=s.dtrain,
                              outlet=s.outlet, hydro_cap=cap, dt=s.dt, manning=s.manning))
    log("serving: all submitted")
    while True:
        try:
            srv.wait(timeout=30.0)
            break
        except pkg.Sw2dError:
            st = srv.stats()
            log(f"serving: {st['completed']} done, {st['running']} running, {st['queued']} queued")
    wall = time.perf_counter() - t0
    st = srv.stats()
    res = [srv.poll(i) for i in ids]
    iters = [r["iterations"] for r in res]
    vol_err = max(abs(r["balance"]["error"]) / max(r["balance"]["volume"], 1e-30) for r in res)
    t_plain = plain_time(pkg, [(s.z.shape[0], k) for s, k in zip(specs, iters)], plain_rates)
    bytes_alg = st["cell_updates"] * bytes_per_cell("fp32", False)
    achieved = bytes_alg / (st["kernel_ms"] / 1e3) / 1e9 if st["kernel_ms"] > 0 else None
    srv.close()
    sides = [s.z.shape[0] for s in specs]
    return {
        "workload": f"{n} requests, Poisson {rate}/s over {seconds} s, sides "
                    f"{sorted(set(sides))} (p 0.4/0.3/0.15/0.1/0.05), {hours} h storms, dx 10 m, "
                    "fixed dt 0.5 s, FP32 state / FP64 balance, admission by HBM footprint and "
                    f"at most {active_cells} running cells",
        "requests": n,
        "completed": st["completed"],
        "failed": st["failed"],
        "cells_mean": float(np.mean([s.cells for s in specs])),
        "wall_s": wall,
        "value": st["cell_updates"] / wall,
        "unit": "cell-updates/s",
        "busy_s": st["busy_seconds"],
        "p50_time_to_hydrograph_s": st["p50_latency"],
        "p99_time_to_hydrograph_s": st["p99_latency"],
        "max_time_to_hydrograph_s": st["max_latency"],
        "p50_queue_s": st["p50_wait"],
        "p99_queue_s": st["p99_wait"],
        "pool_peak_bytes": st["pool_peak"],
        "tile_rows": tile if tile > 0 else "adaptive",
        "retiles": st["retiles"],
        "roofline": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "bytes_per_cell_update": bytes_per_cell("fp32", False)},
        "max_volume_error_rel": vol_err,
        "offered_cell_updates_per_s": float(sum(s.cells * k for s, k in zip(specs, iters))
                                            / max(arrivals[-1], 1e-9)),
        "plain_fp64_one_at_a_time_s": t_plain,
        "scenario_throughput_vs_plain_fp64": t_plain / wall,
    }


def config0_leg(pkg, cases):
    """The reference's own CPU-runnable case (256x256, dx 10 m, 1 h of rain and
    recession, fixed dt 0.5 s, 3600 iterations), the whole run on the GPU:
    the FP64 fused kernel (graph replays).  Returns the leg's figures and
    (spec, iterations, final depth) for the CPU baseline leg, which times the
    FP64 oracle on the same run and compares the final states."""
    spec = cases.config0(outlet=True)
    R, C = spec.z.shape
    iters = int(round(spec.t_end / (2 * spec.dt)))
    c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
    c.set_rain(spec.rain, spec.dtrain)
    c.set_outlet(*spec.outlet)
    c.step(6)
    c.sync()
    c.close()   # the first case paid the one-off setup (graph capture, module load)
    c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
    c.set_rain(spec.rain, spec.dtrain)
    c.set_outlet(*spec.outlet)
    c.sync()
    t0 = time.perf_counter()
    c.step(iters)
    c.sync()
    gpu_s = time.perf_counter() - t0
    h_gpu = c.field("h")
    c.close()
    return {
        "workload": "config 0: 256x256, dx 10 m, 1 % slope + fBm 2 m, n 0.03, rain 50 mm/h for "
                    "30 min then 30 min recession, fixed dt 0.5 s, outlet weir; 3600 iterations",
        "iterations": iters,
        "gpu_fp64_s": gpu_s,
        "gpu_cell_updates_per_s": R * C * iters / gpu_s,
    }, (spec, iters, h_gpu)


def output_leg(pkg, spec, iters, every, precision="fp64"):
    """The output path: every `every` iterations the five fields the driver
    writes (depth, the two cell-centre velocities, the two coupling
    velocities), taken (a) with blocking copies into host arrays
    (sw2d_get_field, the reference's output block) and (b) as double-buffered
    snapshots that a writer thread waits for (views of the pinned buffers)
    and releases while the run goes on.  Wall time of `iters` iterations each
    way beside the same iterations with no output: the cost of an output to
    the stepping thread, file formatting not included (the driver does that
    on its writer thread in both cases).  Snapshot buffers and the derive
    scratch are made before the timed part."""
    import threading
    import queue
    fields = ("h", "uu", "vv", "uua", "vva")
    R, C = spec.z.shape
    wall = {}
    for mode in ("none", "blocking", "snapshot"):
        c = make_case(pkg, spec, precision, "fused", 0)
        c.step(every)
        if mode == "blocking":
            bufs = {n: np.zeros((R, C)) for n in fields}   # allocated once, as the driver does
            c.field("h", bufs["h"])
        q, th = queue.Queue(), None
        if mode == "snapshot":
            ts = [c.snapshot(fields), c.snapshot(fields)]
            for t in ts:
                c.snapshot_release(t)

            def writer():
                while (t := q.get()) is not None:
                    for n in fields:
                        c.snapshot_field(t, n, copy=False)
                    c.snapshot_release(t)
            th = threading.Thread(target=writer)
            th.start()
        c.sync()
        t0 = time.perf_counter()
        for _ in range(iters // every):
            c.step(every)
            if mode == "blocking":
                for n in fields:
                    c.field(n, bufs[n])
            elif mode == "snapshot":
                q.put(c.snapshot(fields, timeout=600.0))
        if th is not None:
            q.put(None)
            th.join()
        c.sync()
        wall[mode] = time.perf_counter() - t0
        c.close()
    n_out = iters // every
    return {
        "workload": f"{R}x{C} {precision} fused, {iters} iterations, an output every {every} "
                    f"({n_out} outputs of {len(fields)} fields, {len(fields) * R * C * 8 / 1e6:.0f} MB each)",
        "wall_s": wall,
        "output_cost_ms": {m: 1e3 * (wall[m] - wall["none"]) / n_out for m in ("blocking", "snapshot")},
    }


def pond_break_leg(pkg, cases, side, sample=30.0):
    """The pond-break case (square reservoir 4 m deep released onto a dry
    fBm floodplain, dx 2 m, n 0.035, adaptive step, 30 min) in the FP32 form
    beside the FP64 form, which equals the CPU FP64 oracle value for value:
    max|dh|, L2(h), first-arrival time of h > 1 cm (sampled every `sample` s)
    and the FP32 ledger's volume error."""
    spec = cases.config2(rows=side, cols=side)
    res = {}
    for prec in ("fp64", "fp32"):
        c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                     precision=prec, adaptive=True, courant=spec.courant)
        arrival = np.full(spec.z.shape, np.inf)
        t, iters, t0 = 0.0, 0, time.perf_counter()
        while t < spec.t_end - 1e-9:
            t = min(t + sample, spec.t_end)
            iters += c.run_until(t)
            h = c.field("h")
            arrival = np.where((h > 0.01) & np.isinf(arrival), c.time, arrival)
        b = c.balance()
        res[prec] = dict(h=h, arrival=arrival, time=c.time, iters=iters,
                         wall=time.perf_counter() - t0, vol_err=b["error"] / b["volume"])
        c.close()
    dh = res["fp32"]["h"] - res["fp64"]["h"]
    a64, a32 = res["fp64"]["arrival"], res["fp32"]["arrival"]
    both = np.isfinite(a64) & np.isfinite(a32)
    return {
        "workload": f"pond break {side}x{side}, dx 2 m, 400 m square reservoir at 4 m, n 0.035, "
                    f"adaptive step, {spec.t_end:.0f} s",
        "max_abs_dh_m": float(np.max(np.abs(dh))),
        "l2_dh_m": float(np.sqrt(np.mean(dh * dh))),
        "max_arrival_error_s": float(np.max(np.abs(a32[both] - a64[both]))) if both.any() else 0.0,
        "arrival_sampling_s": sample,
        "cells_reached_fp64": int(np.isfinite(a64).sum()),
        "cells_reached_fp32": int(np.isfinite(a32).sum()),
        "volume_error_rel_fp32": res["fp32"]["vol_err"],
        "volume_error_rel_fp64": res["fp64"]["vol_err"],
        "iterations": {p: res[p]["iters"] for p in res},
        "simulated_s": {p: res[p]["time"] for p in res},
        "wall_s": {p: res[p]["wall"] for p in res},
    }


def cpu_baseline(spec, seconds, threads, config0=None):
    """Time the FP64 oracle (OpenMP over rows) on the same case for about
    `seconds` of CPU work.  With config0 = (spec, iterations, final GPU depth)
    it also times the oracle's whole run of that case and checks the GPU's
    final depths against it value for value."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    c0 = None
    if config0 is not None:
        spec0, iters0, h_gpu = config0
        o0 = po.Oracle(spec0.z, spec0.h0, spec0.cellsize, dt=spec0.dt, manning=spec0.manning,
                       threads=threads)
        o0.set_rain(spec0.rain, spec0.dtrain)
        o0.set_outlet(*spec0.outlet)
        t0 = time.perf_counter()
        o0.step(iters0)
        cpu_s = time.perf_counter() - t0
        c0 = {"cpu_oracle_s": cpu_s, "cpu_threads": threads,
              "cpu_cell_updates_per_s": spec0.cells * iters0 / cpu_s,
              "final_h_equal": bool(np.array_equal(h_gpu, o0.field("h")))}
        del o0
    o = po.Oracle(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                  threads=threads)
    if spec.rain is not None:
        o.set_rain(spec.rain, spec.dtrain)
    if spec.nclass is not None:
        o.set_manning_map(spec.ntable[spec.nclass])
    if spec.storm is not None:
        s = spec.storm
        o.set_storm(s["peak"], s["x0"], s["y0"], s["vx"], s["vy"], s["sigma"], s["t0"], s["t1"])
    if spec.outlet is not None:
        o.set_outlet(*spec.outlet)
    if spec.lake_mask is not None:
        rec = po.couple_bc(spec.lake_mask, *(np.zeros(spec.z.shape),) * 3, spec.cellsize,
                           spec.cellsize)
        cells = np.concatenate([rec[k]["cell"] for k in ("south", "north", "east", "west")])
        kinds = np.concatenate([np.full(len(rec[k]["cell"]), i)
                                for i, k in enumerate(("south", "north", "east", "west"))])
        o.set_bc(cells, kinds)
    if spec.adaptive:
        o.set_adaptive(spec.courant, 2.0 * spec.dt)
    o.step(1)  # first touch of the arrays
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    out = dict(value=spec.cells * n / el, unit="cell-updates/s", cores=threads, kind="port",
               sample=f"{n} iterations of the same {spec.z.shape[0]}x{spec.z.shape[1]} case "
                      f"({el:.1f} s), FP64 oracle, OpenMP {threads} threads",
               cpu_model=cpu_model(), host_cpus=os.cpu_count())
    return out, c0


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--kernel", default="fused", choices=["fused", "plain"])
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="rows per tile of the fused kernel (0 = measured by sw2d_autotune)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the run, stream this many MB through the library's probe kernel "
                         "(reads and writes n doubles: a copy of known size, to scale rocprofv3 "
                         "FETCH/WRITE_SIZE)")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="iterations of the plain FP64 kernel sequence for the comparison "
                         "line (0 = skip)")
    ap.add_argument("--fp32-steps", type=int, default=3000,
                    help="iterations of the FP32 form on the same case (0 = skip)")
    ap.add_argument("--batch-scenarios", type=int, default=256,
                    help="scenarios of the batched ensemble leg (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="length of the Poisson arrival window of the serving leg (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-hours", type=float, default=2.0)
    ap.add_argument("--sustained-rate", type=float, default=4.0,
                    help="second serving session at this arrival rate (0 = skip)")
    ap.add_argument("--sustained-seconds", type=float, default=10.0)
    ap.add_argument("--serve-tile-rows", type=int, default=0,
                    help="server tile height (0 = adaptive: retiled to the running set)")
    ap.add_argument("--serve-active-cells", type=int, default=32 << 20,
                    help="server admission cap on running cells (0 = HBM footprint only)")
    ap.add_argument("--pond-break", type=int, default=4096,
                    help="side of the pond-break precision study (FP32 form vs FP64; 0 = skip)")
    ap.add_argument("--output-every", type=int, default=100,
                    help="output-path leg: iterations between outputs (0 = skip the leg)")
    ap.add_argument("--output-iters", type=int, default=1000)
    ap.add_argument("--retune", type=int, default=1000,
                    help="drop-in legs: re-time the tile height every N iterations (0 = once)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config0", dest="config0", action="store_false",
                    help="skip the config-0 leg (whole 256x256 run, GPU vs CPU oracle)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 "
                                   "(sharding one grid or a batch over GPUs is not built)"}))
        sys.exit(2)

    pkg = load_package()
    cases = __import__("sw2d_efdc_amd.cases", fromlist=["config1"])
    spec = cases.config1(rows=args.size, cols=args.size)
    R, C = spec.z.shape
    cells = R * C

    # measured bandwidth ceiling beside the 8 TB/s vendor peak: a streaming
    # device copy of 1 GiB (read + write), 20 times
    copy_gbs = pkg.sw2d.stream_copy_gbs(1 << 30, 20)
    log(f"stream copy ceiling: {copy_gbs:.0f} GB/s")
    log(f"{R}x{C} case ready; timing {args.precision} {args.kernel}")
    case, r = timed_run(pkg, spec, args.precision, args.kernel, args.tile_rows, args.warmup,
                        args.steps, args.retune)
    log(f"{args.precision}: {r['value']:.3e} cell-updates/s, tile rows {r['tile_rows']}")
    value, ms_per_step, tile_rows = r["value"], r["ms_per_step"], r["tile_rows"]
    t_sim, dt2, bal, bpc, achieved, avg_s, ws_mb = (r["t_sim"], r["dt2"], r["balance"], r["bpc"],
                                                     r["achieved"], r["avg_s"], r["ws_mb"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 state / f64 balance",
        "data": "synthetic (seeded DEM, storm and lake generated by sw2d-efdc_amd/cases.py)",
        "config": {
            "workload": f"drop-in real-size case: {R}x{C} grid, dx=5 m, slope+3 channels+fBm 3 m"
                        "+prefilled lake pit (EFDC inflow rim), n 0.02/0.08 map, moving Gaussian "
                        "storm 80 mm/h, CFL-adaptive step (Courant 0.5)",
            "grid": [R, C],
            "kernel": args.kernel,
            "precision": args.precision,
            "tile_rows": tile_rows,
            "retune_every": args.retune if args.kernel == "fused" else 0,
            "tile_rows_timed": [[n, H] for n, H in r["heights"]],
            "simulated_s_after_run": t_sim,
            "dt2_last_s": dt2,
            "device_bytes": r["device_bytes"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "stream_copy_GBs": copy_gbs,
            "frac_of_stream_copy": achieved / copy_gbs,
            "traffic": measured_traffic_mix(args.precision, r["heights"], [R, C]),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, scaled by the 256 MiB probe "
                              "(fetch x2.00, write x1.00); profiles/fused_step_2048_dropin.md",
            "bytes_per_cell_update": bpc,
            "kernel_avg_us": avg_s * 1e6,
            "working_set_MB": ws_mb,
        },
        "volume_error_rel": bal["error"] / max(bal["volume"], 1e-30),
    }

    h64 = case.field("h") if args.precision == "fp64" else None
    case.close()
    del case
    if args.fp32_steps > 0 and args.kernel == "fused" and args.precision == "fp64":
        c32, r32 = timed_run(pkg, spec, "fp32", "fused", 0, args.warmup, args.fp32_steps,
                             args.retune)
        cmp32 = None
        if args.fp32_steps == args.steps:
            dh = c32.field("h") - h64
            cmp32 = {"max_abs_dh_m": float(np.max(np.abs(dh))),
                     "l2_dh_m": float(np.sqrt(np.mean(dh * dh))),
                     "t_fp64_s": r["t_sim"], "t_fp32_s": r32["t_sim"]}
        c32.close()
        log(f"fp32: {r32['value']:.3e} cell-updates/s, tile rows {r32['tile_rows']}")
        out["fp32_form"] = {
            "value": r32["value"], "unit": "cell-updates/s", "tile_rows": r32["tile_rows"],
            "tile_rows_timed": [[n, H] for n, H in r32["heights"]],
            "steps": args.fp32_steps, "ms_per_step": r32["ms_per_step"],
            "roofline": {"achieved": r32["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": r32["achieved"] / HBM_PEAK_GBS,
                         "traffic": measured_traffic_mix("fp32", r32["heights"], [R, C]),
                         "bytes_per_cell_update": r32["bpc"],
                         "kernel_avg_us": r32["avg_s"] * 1e6},
            "volume_error_rel": r32["balance"]["error"] / max(r32["balance"]["volume"], 1e-30),
            "vs_fp64_same_iterations": cmp32,
        }
    if args.plain_steps > 0 and args.kernel == "fused" and args.precision == "fp64":
        plain = make_case(pkg, spec, "fp64", "plain", 0)
        plain.step(5)
        plain.sync()
        t0 = time.perf_counter()
        plain.step(args.plain_steps)
        plain.sync()
        pw = time.perf_counter() - t0
        pv = cells * args.plain_steps / pw
        log(f"plain fp64: {pv:.3e} cell-updates/s")
        out["plain_fp64"] = {"value": pv, "unit": "cell-updates/s",
                             "steps": args.plain_steps,
                             "fused_speedup": value / pv}
        del plain

    if args.output_every > 0:
        out["output_path"] = output_leg(pkg, spec, args.output_iters, args.output_every)
        oc = out["output_path"]["output_cost_ms"]
        log(f"outputs: {oc['blocking']:.1f} ms each blocking, {oc['snapshot']:.1f} ms as snapshots")

    if args.pond_break > 0:
        out["precision_study"] = pond_break_leg(pkg, cases, args.pond_break)
        log(f"pond break: max|dh| {out['precision_study']['max_abs_dh_m']:.2e} m")

    plain_rates = {}
    if args.batch_scenarios > 0:
        out["batch"] = batch_leg(pkg, cases, args.batch_scenarios, plain_rates)
        log(f"batch: {out['batch']['value']:.3e} cell-updates/s")
    if args.serve_seconds > 0:
        out["serving"] = serve_leg(pkg, cases, args.serve_seconds, args.serve_rate,
                                   args.serve_hours, plain_rates, args.serve_active_cells,
                                   args.serve_tile_rows)
        log(f"serving: p50 {out['serving']['p50_time_to_hydrograph_s']:.2f} s, "
            f"p99 {out['serving']['p99_time_to_hydrograph_s']:.2f} s")
    if args.serve_seconds > 0 and args.sustained_rate > 0:
        # the same request mix at a rate one GPU keeps up with (~half its
        # capacity): latency without a standing queue
        out["serving_sustained"] = serve_leg(pkg, cases, args.sustained_seconds,
                                             args.sustained_rate, args.serve_hours, plain_rates,
                                             args.serve_active_cells, args.serve_tile_rows)
        log(f"serving at {args.sustained_rate}/s: p50 "
            f"{out['serving_sustained']['p50_time_to_hydrograph_s']:.2f} s, p99 "
            f"{out['serving_sustained']['p99_time_to_hydrograph_s']:.2f} s")

    c0_run = None
    if not args.no_cpu and args.config0:
        out["config0"], c0_run = config0_leg(pkg, cases)
        log(f"config 0: GPU {out['config0']['gpu_fp64_s']:.3f} s")
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads, {args.cpu_seconds} s")
        out["cpu_baseline"], c0_cpu = cpu_baseline(spec, args.cpu_seconds, threads, c0_run)
        if c0_cpu is not None:
            out["config0"].update(c0_cpu)
            out["config0"]["speedup"] = c0_cpu["cpu_oracle_s"] / out["config0"]["gpu_fp64_s"]
            log(f"config 0: CPU oracle {c0_cpu['cpu_oracle_s']:.2f} s ({threads} threads), "
                f"final h equal: {c0_cpu['final_h_equal']}")
    else:
        out["cpu_baseline"] = None
    if args.calibrate_copy_mb > 0:
        # the library's lane-move probe reads n doubles and writes n doubles
        # (sw2d_device_math fn 6): a streaming kernel of known size
        n = args.calibrate_copy_mb * (1 << 20) // 8
        pkg.sw2d.device_math(6, np.ones(n))
        out["copy_calibration_bytes"] = {"kernel": "math_probe", "read": 8 * n, "write": 8 * n}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
