#!/usr/bin/env python3
"""Benchmark of the SW2D time step on one MI355X.

Workload (default): the 2048 x 2048 "plain FP64 GPU drop-in at real size"
case -- dx = 5 m, 0.5 % slope + 3 Gaussian-valley channels + fBm 3 m +
prefilled Gaussian lake pit whose rim cells are the EFDC inflow records,
Manning 0.02 (channels) / 0.08 (hillslopes), moving Gaussian storm (sigma
1.5 km, 80 mm/h, 5 m/s, 1 h), CFL-adaptive step (Courant 0.5) -- run through
the fused FP64 kernel.  One step = one loop iteration of the reference
(flux, continuity, treat_error, continuity, derived fields) over all cells;
the metric is cell-updates per second with the state resident in HBM.

Printed: one JSON line with the value, the roofline of the step kernel
(algorithmic bytes per launch / average launch time from HIP events on the
library's stream) and the CPU baseline (the FP64 oracle, OpenMP, on a bounded
sample of the same workload).
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


def cpu_baseline(spec, seconds, threads):
    """Time the FP64 oracle (OpenMP over rows) on the same case for about
    `seconds` of CPU work."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
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
    return dict(value=spec.cells * n / el, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"{n} iterations of the same {spec.z.shape[0]}x{spec.z.shape[1]} case "
                       f"({el:.1f} s), FP64 oracle, OpenMP {threads} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--kernel", default="fused", choices=["fused", "plain"])
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="rows per tile of the fused kernel (0 = measured by sw2d_autotune)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the run, copy this many MB device-to-device with torch (a plain "
                         "copy kernel of known size, to scale rocprofv3 FETCH/WRITE_SIZE)")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="iterations of the plain FP64 kernel sequence for the comparison "
                         "line (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        sys.exit(2)

    pkg = load_package()
    cases = __import__("sw2d_efdc_amd.cases", fromlist=["config1"])
    spec = cases.config1(rows=args.size, cols=args.size)
    R, C = spec.z.shape
    cells = R * C

    case = make_case(pkg, spec, args.precision, args.kernel, args.tile_rows)
    tile_rows = args.tile_rows
    if args.kernel == "fused" and tile_rows == 0:
        tile_rows = case.autotune(10)
    case.step(args.warmup)
    case.sync()

    # timed region: K iterations, HIP events around every launch on the library stream
    case.timing(True)
    case.sync()
    t0 = time.perf_counter()
    case.step(args.steps)
    case.sync()
    t1 = time.perf_counter()
    kms, launches = case.kernel_time()
    case.timing(False)
    wall = t1 - t0
    value = cells * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    t_sim, mstep, dt2 = case.clock()
    bal = case.balance()

    bpc = bytes_per_cell(args.precision, spec.nclass is not None)
    per_launch = bpc * cells
    avg_s = (kms / 1e3) / max(launches, 1)
    achieved = per_launch / avg_s / 1e9
    ws_mb = case.device_bytes() / 1e6

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
            "simulated_s_after_run": t_sim,
            "dt2_last_s": dt2,
            "device_bytes": case.device_bytes(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_cell_update": bpc,
            "kernel_avg_us": avg_s * 1e6,
            "working_set_MB": ws_mb,
        },
        "volume_error_rel": bal["error"] / max(bal["volume"], 1e-30),
    }

    if args.plain_steps > 0 and args.kernel == "fused" and args.precision == "fp64":
        del case
        plain = make_case(pkg, spec, "fp64", "plain", 0)
        plain.step(5)
        plain.sync()
        t0 = time.perf_counter()
        plain.step(args.plain_steps)
        plain.sync()
        pw = time.perf_counter() - t0
        pv = cells * args.plain_steps / pw
        out["plain_fp64"] = {"value": pv, "unit": "cell-updates/s",
                             "steps": args.plain_steps,
                             "fused_speedup": value / pv}
        del plain

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_seconds, threads)
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
