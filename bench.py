#!/usr/bin/env python3
"""Benchmark of the SW2D time step on one MI355X.

Headline (`value`): the serving workload -- Poisson arrivals at 20 requests/s,
grid sides 256..4096 (p 0.4/0.3/0.15/0.1/0.05), dx 10 m, seeded fBm DEMs,
2 h block storms, fixed dt 0.5 s -- served by the continuous-batching server
(HBM-footprint admission, state resident on the device) in the FP32 form
(FP32 state and fluxes, FP64 dt, volume ledger and hydrograph).  `value` is
the cell-updates per second the server delivered over the whole session
(first arrival to last hydrograph); p50/p99 time-to-hydrograph are reported
absolute and as a fraction of each request's HBM-roofline time.  The arrival
window defaults to 30 s of the 600 s the workload names (the offered load is
about 2.5x one GPU's capacity, so the full window takes ~25 min to drain);
`config.workload` states the window.

One cell-update is one grid cell advanced by one loop iteration of the
reference (flux, continuity, treat_error, continuity, derived fields).

Also measured (same JSON line), each with max|dh|, L2(h), peak-outflow and
volume error against FP64 for every reduced-precision result:
- the drop-in case (2048^2 storm, lake pit, Manning map, adaptive step) in
  FP64, FP32 and the mixed form, timed on a wetted window (30 min into the
  storm, whatever --warmup is) for --steps iterations; the plain (unfused)
  FP64 kernel sequence on the same case;
- the 256-scenario ensemble (128..1024 grids, 1 h storms) as one batch;
- the plain FP64 path one scenario at a time on measured samples of the
  served and batched scenarios (full storms), for the throughput ratios;
- the pond break (4096^2, 30 min), the reference's own 256^2 case (whole run,
  GPU against the CPU oracle), the output path, a sustainable-rate serving
  session, and the CPU baseline (the FP64 oracle, OpenMP, on a bounded sample
  of the served workload).
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
SERVE_TRAFFIC = ROOT / "profiles" / "serving_traffic.json"
SERVE_ISSUE = ROOT / "profiles" / "serving_issue.json"
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


def bytes_per_cell(precision, has_class_map=False):
    """Compulsory HBM bytes of one cell-update of the fused kernel: read the
    depth h, the fluxes um, vn and the bed (the FP64 z, or in the FP32 form
    the cell's two FP32 bed steps: 8 B either way), write h, um, vn (+1 byte
    Manning class with a class map): 57 B in FP64, 33 B in the FP32 form,
    41 B in the mixed form (FP64 depth, FP32 fluxes)."""
    hb = {"fp64": 8, "fp32": 4, "mixed": 8}[precision]
    fb = {"fp64": 8, "fp32": 4, "mixed": 4}[precision]
    return 2 * hb + 4 * fb + 8 + (1 if has_class_map else 0)


def errors_vs_fp64(h, h64, q_peak, q_peak64, bal):
    """max|dh|, L2(h), relative peak-outflow error and relative volume error."""
    dh = np.asarray(h) - np.asarray(h64)
    return {"max_abs_dh_m": float(np.max(np.abs(dh))),
            "l2_dh_m": float(np.sqrt(np.mean(dh * dh))),
            "peak_outflow_err_rel": (abs(q_peak - q_peak64) / q_peak64) if q_peak64 > 0 else 0.0,
            "volume_error_rel": abs(bal["error"]) / max(bal["volume"], 1e-30)}


def worst(errs):
    """Largest of each error over a list of errors_vs_fp64 records."""
    if not errs:
        return None
    return {k: max(e[k] for e in errs) for k in errs[0]}


def spread_summary(errs, noise_errs):
    """Errors of a reduced form beside the FP64 run's own spread under
    sub-survey DEM noise (the same scenarios, bed + 1 um noise): worst and
    median of each, members above 1 mm, and the members whose max|dh| is
    within twice the FP64 noise spread (or below 1 mm)."""
    md = np.array([e["max_abs_dh_m"] for e in errs])
    out = {"worst": worst(errs), "median_max_abs_dh_m": float(np.median(md)),
           "members": len(errs), "members_above_1mm": int((md > 1e-3).sum()),
           "members_above_1cm": int((md > 1e-2).sum())}
    if noise_errs:
        nd = np.array([e["max_abs_dh_m"] for e in noise_errs])
        out["fp64_dem_noise_1um"] = {"worst": worst(noise_errs),
                                     "median_max_abs_dh_m": float(np.median(nd)),
                                     "members_above_1mm": int((nd > 1e-3).sum()),
                                     "members_above_1cm": int((nd > 1e-2).sum())}
        out["members_within_fp64_noise_spread"] = int(((md <= 2 * nd) | (md <= 1e-3)).sum())
    return out


def make_case(pkg, spec, precision, kernel, tile_rows=0):
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


def batch_run(pkg, specs, precision, tile_rows=0, tune_after=300, dem_noise=0.0):
    """The scenarios as one batch to their end times: members, wall seconds,
    kernel (ms, launches), tile height (tuned after `tune_after` iterations when
    tile_rows is 0).  dem_noise > 0 adds seeded N(0, dem_noise^2) metres to
    every bed elevation (the reference's sensitivity yardstick)."""
    need = sum(pkg.footprint(s.z.shape[0], s.z.shape[1], precision=precision,
                             batch_tile_rows=8, hydro_cap=int(s.t_end / (2 * s.dt)) + 16)
               for s in specs)
    b = pkg.Batch(precision=precision, pool_bytes=int(need * 1.1) + (64 << 20),
                  tile_rows=tile_rows)
    members = []
    rng = np.random.default_rng(11)
    for s in specs:
        z = s.z + dem_noise * rng.standard_normal(s.z.shape) if dem_noise > 0 else s.z
        m = b.case(z, s.h0, s.cellsize, dt=s.dt, manning=s.manning)
        m.set_rain(s.rain, s.dtrain)
        m.set_outlet(*s.outlet)
        m.set_hydrograph(int(s.t_end / (2 * s.dt)) + 16)
        b.attach(m, t_end=s.t_end)
        members.append(m)
    b.timing(True)
    t0 = time.perf_counter()
    H = tile_rows
    b.tuning = None
    if tile_rows == 0 and tune_after > 0:
        # the tile height and the launch schedule are tuned once the rain has
        # wetted the slopes (on the dry initial state the flux work is
        # skipped); trials are timed in
        b.step(tune_after)
        H = b.autotune(8)
        order, waves = b.schedule()
        b.tuning = {"candidates": [{"tile_rows": int(r[0]), "order": int(r[1]),
                                    "waves_per_simd_cap": int(r[2]), "ms_per_launch": float(r[3])}
                                   for r in b.tune_report()],
                    "pick": {"tile_rows": H, "order": order, "waves_per_simd_cap": waves}}
    while b.step(60):
        pass
    wall = time.perf_counter() - t0
    kms, launches = b.kernel_time()
    return b, members, wall, kms, launches, H


def member_errors(members, members64):
    """Errors of batch members against their FP64 twins.  Depths are read with
    sw2d_get_field_noalloc: a batch member's arrays live in the batch pool,
    which has no room for a derive buffer of a large grid."""
    out = []
    for m, m64 in zip(members, members64):
        b, b64 = m.balance(), m64.balance()
        out.append(errors_vs_fp64(m.field_noalloc("h"), m64.field_noalloc("h"), b["q_peak"],
                                  b64["q_peak"], b))
    return out


def plain_seconds(pkg, spec):
    """Wall seconds of the plain FP64 kernel sequence running one scenario over
    its whole storm, alone (the reference's structure: one grid per process)."""
    c = make_case(pkg, spec, "fp64", "plain")
    c.sync()
    t0 = time.perf_counter()
    c.run_until(spec.t_end)
    c.sync()
    s = time.perf_counter() - t0
    c.close()
    return s


def dropin_legs(pkg, cases, size, steps, warmup, wet_time, plain_steps):
    """The drop-in case in FP64, FP32 and the mixed form.  FP64 runs to
    `wet_time` of the storm (n_wet iterations; the reduced forms run the same
    n_wet iterations), then `warmup` more, then the tile height is tuned
    there and exactly `steps` iterations are timed with HIP events around
    every launch (graph replay) on the library's stream.  The reduced forms'
    depths are compared with FP64's after the same iterations."""
    spec = cases.config1(rows=size, cols=size)
    R, C = spec.z.shape
    out, h64, q64, n_wet = {}, None, None, None
    for prec in ("fp64", "fp32", "mixed"):
        c = make_case(pkg, spec, prec, "fused")
        c.set_hydrograph(200000)
        if n_wet is None:
            n_wet = c.run_until(wet_time)
        else:
            c.step(n_wet)
        c.step(warmup)
        H = c.autotune(10)
        t_start = c.time
        c.sync()
        c.timing(True)
        t0 = time.perf_counter()
        c.step(steps)
        c.sync()
        wall = time.perf_counter() - t0
        kms, launches = c.kernel_time()
        c.timing(False)
        bpc = bytes_per_cell(prec, True)
        avg_s = (kms / 1e3) / max(launches, 1)
        bal = c.balance()
        h = c.field("h")
        r = {"value": R * C * steps / wall, "unit": "cell-updates/s", "ms_per_step": 1e3 * wall / steps,
             "tile_rows": H, "iterations_before_timing": n_wet + warmup,
             "simulated_window_s": [t_start, c.time],
             "roofline": {"bound": "hbm", "achieved": bpc * R * C / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": bpc * R * C / avg_s / 1e9 / HBM_PEAK_GBS,
                          "bytes_per_cell_update": bpc, "kernel_avg_us": avg_s * 1e6,
                          "launches_timed": launches},
             "volume_error_rel": abs(bal["error"]) / max(bal["volume"], 1e-30),
             "device_bytes": c.device_bytes()}
        if prec == "fp64":
            h64, q64 = h, bal["q_peak"]
        else:
            r["vs_fp64_same_iterations"] = errors_vs_fp64(h, h64, bal["q_peak"], q64, bal)
        out[prec] = r
        total_iters = n_wet + warmup + steps
        log(f"drop-in {prec}: {r['value']:.3e} cell-updates/s, H {H}, kernel "
            f"{avg_s * 1e6:.1f} us, window t = {t_start:.0f}..{c.time:.0f} s")
        c.close()
    # the FP64 run's own spread: the same iterations with 1 um of seeded noise
    # on the bed (sub-survey), against the unperturbed FP64 run
    rng = np.random.default_rng(11)
    spec_n = cases.config1(rows=size, cols=size)
    spec_n.z = spec.z + 1e-6 * rng.standard_normal(spec.z.shape)
    c = make_case(pkg, spec_n, "fp64", "fused")
    c.set_hydrograph(200000)
    c.step(total_iters)
    bal = c.balance()
    out["fp64_dem_noise_1um"] = errors_vs_fp64(c.field("h"), h64, bal["q_peak"], q64, bal)
    c.close()
    log(f"drop-in fp64 under 1 um DEM noise: max|dh| {out['fp64_dem_noise_1um']['max_abs_dh_m']:.2e} m")
    if plain_steps > 0:
        p = make_case(pkg, spec, "fp64", "plain")
        p.step(n_wet)
        p.sync()
        t0 = time.perf_counter()
        p.step(plain_steps)
        p.sync()
        pv = R * C * plain_steps / (time.perf_counter() - t0)
        p.close()
        out["plain_fp64"] = {"value": pv, "unit": "cell-updates/s", "steps": plain_steps,
                             "fused_fp64_speedup": out["fp64"]["value"] / pv}
        log(f"drop-in plain fp64: {pv:.3e} cell-updates/s")
    out["workload"] = (f"drop-in real-size case: {R}x{C} grid, dx 5 m, slope + 3 channels + fBm 3 m "
                       "+ prefilled lake pit (EFDC inflow rim), n 0.02/0.08 map, moving Gaussian "
                       f"storm 80 mm/h, CFL-adaptive step (Courant 0.5); timed {steps} iterations "
                       f"from t = {wet_time:.0f} s of the storm")
    return out


def batch_leg(pkg, cases, n, plain_sample):
    """The ensemble: n scenarios (128..1024, 1 h, fixed dt 0.5 s) in one batch in
    the FP32 form, with the same batch in FP64 (the errors) and in the mixed
    form; the plain FP64 path on a seeded sample of the scenarios over their
    full storms gives the scenario-throughput ratio."""
    specs = [cases.ensemble_scenario(k) for k in range(n)]
    log(f"batch: {n} scenarios generated")
    b32, m32, wall, kms, launches, H = batch_run(pkg, specs, "fp32")
    tuning = b32.tuning
    iters = [b32.status(m)[1] for m in m32]
    cell_updates = sum(s.cells * k for s, k in zip(specs, iters))
    bpc = bytes_per_cell("fp32")
    achieved = cell_updates * bpc / (kms / 1e3) / 1e9
    log(f"batch fp32: {cell_updates / wall:.3e} cell-updates/s, H {H}")
    b64, m64, wall64, kms64, _, H64 = batch_run(pkg, specs, "fp64")
    log(f"batch fp64: {sum(s.cells * k for s, k in zip(specs, iters)) / wall64:.3e} cell-updates/s")
    e32 = member_errors(m32, m64)
    b32.close()
    bmx, mmx, wallmx, kmsmx, _, Hmx = batch_run(pkg, specs, "mixed")
    emx = member_errors(mmx, m64)
    bmx.close()
    log(f"batch mixed: {cell_updates / wallmx:.3e} cell-updates/s")
    bno, mno, *_ = batch_run(pkg, specs, "fp64", tile_rows=H64, tune_after=0, dem_noise=1e-6)
    eno = member_errors(mno, m64)
    bno.close()
    b64.close()
    log(f"batch: fp32 worst max|dh| {worst(e32)['max_abs_dh_m']:.2e} m, fp64 under 1 um DEM "
        f"noise {worst(eno)['max_abs_dh_m']:.2e} m")
    # plain FP64, one scenario at a time over its full storm, on a seeded
    # sample; the whole ensemble's plain time from the sample's seconds per
    # cell-update at the nearest sampled side (log scale)
    rng = np.random.default_rng(7)
    idx = sorted(rng.choice(n, size=min(plain_sample, n), replace=False).tolist())
    samp = [(specs[i].z.shape[0], plain_seconds(pkg, specs[i]) / (specs[i].cells * iters[i]))
            for i in idx]
    sides = np.array([s for s, _ in samp], dtype=float)
    t_plain = 0.0 if samp else None   # no sample: no plain comparison
    for s, k in zip(specs, iters):
        if not samp:
            break
        j = int(np.argmin(np.abs(np.log(sides / s.z.shape[0]))))
        t_plain += samp[j][1] * s.cells * k
    log(f"batch: plain fp64 sample of {len(idx)} scenarios measured")
    return {
        "workload": f"{n} ensemble scenarios, square grids 128..1024 (log-uniform), dx 10 m, "
                    "slope U(0.2%,3%), fBm U(0.5,5) m, n U(0.02,0.1), block rain U(10,120) mm/h "
                    "for U(15,90) min, free outlet, fixed dt 0.5 s, 1 h each, one launch per "
                    "iteration for all",
        "precision": "f32 state / f64 balance",
        "cells_total": int(sum(s.cells for s in specs)),
        "tile_rows": H,
        "autotune": tuning,
        "iterations": int(max(iters)),
        "wall_s": wall,
        "value": cell_updates / wall,
        "unit": "cell-updates/s",
        "scenarios_per_s": n / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell_update": bpc,
                     "kernel_ms_total": kms, "launches": launches,
                     "kernel_avg_us": 1e3 * kms / max(launches, 1)},
        "vs_fp64": spread_summary(e32, eno),
        "fp64_batch": {"value": cell_updates / wall64, "wall_s": wall64, "tile_rows": H64},
        "mixed_batch": {"value": cell_updates / wallmx, "wall_s": wallmx, "tile_rows": Hmx,
                        "roofline_frac": cell_updates * bytes_per_cell("mixed") / (kmsmx / 1e3) / 1e9
                        / HBM_PEAK_GBS,
                        "vs_fp64": spread_summary(emx, eno)},
        "plain_fp64_sample": {"scenarios": len(idx), "sides": [int(s) for s, _ in samp],
                              "ns_per_cell_update": [1e9 * v for _, v in samp]},
        "plain_fp64_one_at_a_time_s": t_plain,
        "scenario_throughput_vs_plain_fp64": t_plain / wall if t_plain is not None else None,
    }


def serve_leg(pkg, cases, seconds, rate, hours, precision, active_cells, tile, err_sample,
              plain=True, label="serving"):
    """A serving session: Poisson arrivals at `rate`/s for `seconds`, grid
    sides 256..4096 (p 0.4/0.3/0.15/0.1/0.05), `hours`-h storms, in
    `precision`.  With `plain`, the plain FP64 path runs one request of each
    side of the stream over its whole storm (measured), and the stream's
    one-at-a-time plain time is the sum over its requests.  `err_sample`
    requests of the stream are rerun in FP64 and in `precision` as a batch
    (results identical to the server's, value for value) for the errors."""
    n = max(1, int(round(seconds * rate)))
    arrivals, specs = cases.serve_requests(n, rate=rate, t_end=3600.0 * hours)
    log(f"{label}: {n} requests generated; submitting")
    srv = pkg.Server(precision=precision, pool_bytes=64 << 30, tile_rows=tile, slice_iters=60,
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
    log(f"{label}: all submitted")
    while True:
        try:
            srv.wait(timeout=20.0)
            break
        except pkg.Sw2dError:
            st = srv.stats()
            log(f"{label}: {st['completed']} done, {st['running']} running, {st['queued']} queued")
    wall = time.perf_counter() - t0
    st = srv.stats()
    res = [srv.poll(i) for i in ids]
    srv.close()
    bpc = bytes_per_cell(precision)
    iters = [r["iterations"] for r in res]
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    # time the same request's bytes take at the HBM peak, and the latency
    # as a fraction of it
    t_roof = np.array([s.cells * k * bpc / (HBM_PEAK_GBS * 1e9) for s, k in zip(specs, iters)])
    frac = t_roof / lat
    achieved = st["cell_updates"] * bpc / (st["kernel_ms"] / 1e3) / 1e9 if st["kernel_ms"] > 0 else None
    vol_err = max(abs(r["balance"]["error"]) / max(r["balance"]["volume"], 1e-30) for r in res)
    sides = [s.z.shape[0] for s in specs]
    out = {
        "workload": f"{n} requests, Poisson {rate}/s over {seconds} s, sides "
                    f"{sorted(set(sides))} (p 0.4/0.3/0.15/0.1/0.05), {hours} h storms, dx 10 m, "
                    f"fixed dt 0.5 s, {precision}, admission by HBM footprint and at most "
                    f"{active_cells} running cells",
        "requests": n, "completed": st["completed"], "failed": st["failed"],
        "cells_mean": float(np.mean([s.cells for s in specs])),
        "wall_s": wall,
        "value": st["cell_updates"] / wall,
        "unit": "cell-updates/s",
        "busy_s": st["busy_seconds"],
        "launches": st["launches"],
        "p50_time_to_hydrograph_s": st["p50_latency"],
        "p99_time_to_hydrograph_s": st["p99_latency"],
        "max_time_to_hydrograph_s": st["max_latency"],
        # roofline time / latency: 1.0 would be a request served alone at the
        # HBM peak with no queueing; p99 is the 1st percentile (the slowest)
        "p50_time_to_hydrograph_roofline_frac": float(np.percentile(frac, 50)),
        "p99_time_to_hydrograph_roofline_frac": float(np.percentile(frac, 1)),
        "p50_queue_s": st["p50_wait"], "p99_queue_s": st["p99_wait"],
        "pool_peak_bytes": st["pool_peak"],
        "tile_rows": tile if tile > 0 else "adaptive", "retiles": st["retiles"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "bytes_per_cell_update": bpc, "kernel_ms_total": st["kernel_ms"],
                     "kernel_avg_us": 1e3 * st["kernel_ms"] / max(st["launches"], 1)},
        "max_volume_error_rel": vol_err,
        "offered_cell_updates_per_s": float(sum(s.cells * k for s, k in zip(specs, iters))
                                            / max(arrivals[-1], 1e-9)),
    }
    if err_sample > 0:
        # the first request of each side present, then the earliest others
        pick = []
        for sd in sorted(set(sides)):
            pick.append(sides.index(sd))
        for i in range(n):
            if len(pick) >= err_sample:
                break
            if i not in pick:
                pick.append(i)
        pick = sorted(pick[:max(err_sample, len(set(sides)))])
        sub = [specs[i] for i in pick]
        # the served mix's sample as a batch, tuned as the ensemble is: the
        # tuner's candidates and pick for this workload
        bq, mq, *_ = batch_run(pkg, sub, precision)
        b64, m64, *_ = batch_run(pkg, sub, "fp64", tile_rows=32, tune_after=0)
        errs = member_errors(mq, m64)
        bq.close()
        bno, mno, *_ = batch_run(pkg, sub, "fp64", tile_rows=32, tune_after=0, dem_noise=1e-6)
        eno = member_errors(mno, m64)
        bno.close()
        b64.close()
        out["autotune_on_sample"] = bq.tuning
        out["vs_fp64_sample"] = dict(spread_summary(errs, eno), requests=pick,
                                     sides=[sides[i] for i in pick],
                                     max_abs_dh_m=[e["max_abs_dh_m"] for e in errs],
                                     fp64_dem_noise_max_abs_dh_m=[e["max_abs_dh_m"] for e in eno])
        log(f"{label}: errors vs fp64 on {len(pick)} requests: max|dh| "
            f"{out['vs_fp64_sample']['worst']['max_abs_dh_m']:.2e} m (fp64 under 1 um DEM noise "
            f"{worst(eno)['max_abs_dh_m']:.2e} m)")
    if plain:
        per_side = {}
        for sd in sorted(set(sides)):
            i = sides.index(sd)
            per_side[sd] = plain_seconds(pkg, specs[i])
            log(f"{label}: plain fp64 {sd}^2 full storm {per_side[sd]:.2f} s")
        t_plain = sum(per_side[sd] for sd in sides)
        out["plain_fp64_per_side_s"] = {str(k): v for k, v in per_side.items()}
        out["plain_fp64_one_at_a_time_s"] = t_plain
        out["scenario_throughput_vs_plain_fp64"] = t_plain / wall
    return out


def config0_leg(pkg, cases):
    """The reference's own CPU-runnable case (256x256, dx 10 m, 1 h of rain and
    recession, fixed dt 0.5 s, 3600 iterations), the whole run on the GPU in
    FP64.  Returns the leg's figures and (spec, iterations, final depth) for
    the CPU baseline leg, which times the FP64 oracle on the same run and
    compares the final states."""
    spec = cases.config0(outlet=True)
    R, C = spec.z.shape
    iters = int(round(spec.t_end / (2 * spec.dt)))
    for first in (True, False):
        # the first case pays the one-off setup (graph capture, module load)
        c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
        c.set_rain(spec.rain, spec.dtrain)
        c.set_outlet(*spec.outlet)
        c.sync()
        if first:
            c.step(6)
            c.sync()
            c.close()
            continue
        t0 = time.perf_counter()
        c.step(iters)
        c.sync()
        gpu_s = time.perf_counter() - t0
        h_gpu = c.field("h")
        c.close()
    return {
        "workload": "256x256, dx 10 m, 1 % slope + fBm 2 m, n 0.03, rain 50 mm/h for "
                    "30 min then 30 min recession, fixed dt 0.5 s, outlet weir; 3600 iterations",
        "iterations": iters,
        "gpu_fp64_s": gpu_s,
        "gpu_cell_updates_per_s": R * C * iters / gpu_s,
    }, (spec, iters, h_gpu)


def output_leg(pkg, spec, iters, every, precision="fp64"):
    """The output path: every `every` iterations the five fields the driver
    writes, taken (a) with blocking copies into host arrays (sw2d_get_field,
    the reference's output block) and (b) as double-buffered snapshots that a
    writer thread waits for and releases while the run goes on.  Wall time
    of `iters` iterations each way beside the same iterations with no output;
    snapshot buffers and the derive scratch are made before the timed part."""
    import threading
    import queue
    fields = ("h", "uu", "vv", "uua", "vva")
    R, C = spec.z.shape
    wall = {}
    for mode in ("none", "blocking", "snapshot"):
        c = make_case(pkg, spec, precision, "fused", 0)
        c.step(every)
        if mode == "blocking":
            bufs = {n: np.zeros((R, C)) for n in fields}
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
    """The pond break (square reservoir 4 m deep released onto a dry fBm
    floodplain, dx 2 m, n 0.035, adaptive step, 30 min): FP32 and mixed forms
    beside FP64 (which equals the CPU oracle value for value): max|dh|,
    L2(h), first-arrival time of h > 1 cm (sampled every `sample` s) and the
    volume error; and FP64 on the bed plus 1 um of seeded noise, the
    reference's own spread."""
    spec = cases.config2(rows=side, cols=side)
    z_noise = spec.z + 1e-6 * np.random.default_rng(11).standard_normal(spec.z.shape)
    res = {}
    for prec in ("fp64", "fp32", "mixed", "fp64_dem_noise_1um"):
        c = pkg.Case(z_noise if prec.startswith("fp64_") else spec.z, spec.h0, spec.cellsize,
                     dt=spec.dt, manning=spec.manning, precision=prec[:4] if prec.startswith("fp64") else prec,
                     adaptive=True, courant=spec.courant)
        arrival = np.full(spec.z.shape, np.inf)
        t, iters, t0 = 0.0, 0, time.perf_counter()
        while t < spec.t_end - 1e-9:
            t = min(t + sample, spec.t_end)
            iters += c.run_until(t)
            h = c.field("h")
            arrival = np.where((h > 0.01) & np.isinf(arrival), c.time, arrival)
        b = c.balance()
        res[prec] = dict(h=h, arrival=arrival, time=c.time, iters=iters, bal=b,
                         wall=time.perf_counter() - t0)
        c.close()
    out = {"workload": f"pond break {side}x{side}, dx 2 m, 400 m square reservoir at 4 m, n 0.035, "
                       f"adaptive step, {spec.t_end:.0f} s",
           "arrival_sampling_s": sample,
           "iterations": {p: res[p]["iters"] for p in res},
           "simulated_s": {p: res[p]["time"] for p in res},
           "wall_s": {p: res[p]["wall"] for p in res}}
    a64 = res["fp64"]["arrival"]
    for prec in ("fp32", "mixed", "fp64_dem_noise_1um"):
        r = res[prec]
        e = errors_vs_fp64(r["h"], res["fp64"]["h"], 1.0, 1.0, r["bal"])
        del e["peak_outflow_err_rel"]
        a = r["arrival"]
        both = np.isfinite(a64) & np.isfinite(a)
        e["max_arrival_error_s"] = float(np.max(np.abs(a[both] - a64[both]))) if both.any() else 0.0
        e["cells_reached"] = int(np.isfinite(a).sum())
        out[prec] = e
    out["cells_reached_fp64"] = int(np.isfinite(a64).sum())
    out["volume_error_rel_fp64"] = abs(res["fp64"]["bal"]["error"]) / res["fp64"]["bal"]["volume"]
    return out


def cpu_model():
    """The host CPU's model name with its family / model / stepping numbers
    (from /proc/cpuinfo: a virtual machine may report only "AMD EPYC", the
    numbers still tell the generation), or None."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break
                k, _, v = line.partition(":")
                info.setdefault(k.strip(), v.strip())
    except OSError:
        return None
    name = info.get("model name")
    if not name:
        return None
    ids = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
    return f"{name} ({', '.join(ids)})" if ids else name


def cpu_baseline(cases, seconds, threads, hours, config0=None):
    """The FP64 oracle (OpenMP over rows) on a bounded sample of the served
    workload -- the first 1024^2 request of the serving stream, iterations of
    its storm for about `seconds` -- on the host's cores.  With config0 =
    (spec, iterations, final GPU depth) it also times the oracle's whole run of
    the reference's 256^2 case and checks the GPU's final depths against it
    value for value."""
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
    _, specs = cases.serve_requests(64, t_end=3600.0 * hours)
    spec = next(s for s in specs if s.z.shape[0] == 1024)
    o = po.Oracle(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                  threads=threads)
    cases.apply(o, spec)
    o.step(300)  # into the storm (first touch of the arrays too)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    out = dict(value=spec.cells * n / el, unit="cell-updates/s", cores=threads, kind="port",
               sample=f"{n} iterations (after 300) of a 1024x1024 request of the serving stream "
                      f"({el:.1f} s), FP64 oracle, OpenMP {threads} threads",
               cpu_model=cpu_model(), host_cpus=os.cpu_count())
    return out, c0


def serve_traffic(precision):
    """HBM bytes per cell-update of the served kernel, measured by rocprofv3
    FETCH_SIZE / WRITE_SIZE passes over the serving session and committed
    under profiles/ (None when no measurement of this form is recorded)."""
    try:
        t = json.load(open(SERVE_TRAFFIC))
    except (OSError, ValueError):
        return None
    return t.get("bytes_per_cell_update", {}).get(precision)


def serve_issue(precision):
    """VALU wave-instructions per cell-update of the served kernel (rocprofv3
    SQ counters over the serving session, committed under profiles/) and the
    VALU issue peak; None when no measurement of this form is recorded."""
    try:
        t = json.load(open(SERVE_ISSUE))
    except (OSError, ValueError):
        return None
    per = t.get("valu_wave_instructions_per_cell_update", {}).get(precision)
    peak = t.get("valu_issue_peak_wave_instructions_per_s")
    return (per, peak) if per and peak else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed iterations of each drop-in leg")
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed iterations of each drop-in leg after its wetted start")
    ap.add_argument("--size", type=int, default=2048, help="side of the drop-in case")
    ap.add_argument("--wet-time", type=float, default=1800.0,
                    help="drop-in legs: simulated seconds of the storm before timing")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="iterations of the plain FP64 sequence on the drop-in case (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="length of the Poisson arrival window of the headline serving session")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-hours", type=float, default=2.0)
    ap.add_argument("--serve-precision", default="fp32", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--serve-tile-rows", type=int, default=0,
                    help="server tile height (0 = adaptive: retiled to the running set)")
    ap.add_argument("--serve-active-cells", type=int, default=64 << 20,
                    help="server admission cap on running cells (0 = HBM footprint only)")
    ap.add_argument("--serve-err-sample", type=int, default=8,
                    help="requests of the stream rerun in FP64 for the errors (0 = skip)")
    ap.add_argument("--no-serve-plain", dest="serve_plain", action="store_false",
                    help="skip the plain FP64 runs of the serving comparison")
    ap.add_argument("--sustained-rate", type=float, default=4.0,
                    help="second serving session at this arrival rate (0 = skip)")
    ap.add_argument("--sustained-seconds", type=float, default=10.0)
    ap.add_argument("--batch-scenarios", type=int, default=256,
                    help="scenarios of the batched ensemble leg (0 = skip)")
    ap.add_argument("--batch-plain-sample", type=int, default=12)
    ap.add_argument("--pond-break", type=int, default=4096,
                    help="side of the pond-break precision study (0 = skip)")
    ap.add_argument("--output-every", type=int, default=100,
                    help="output-path leg: iterations between outputs (0 = skip the leg)")
    ap.add_argument("--output-iters", type=int, default=1000)
    ap.add_argument("--no-dropin", dest="dropin", action="store_false")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config0", dest="config0", action="store_false",
                    help="skip the whole 256x256 run (GPU vs CPU oracle)")
    ap.add_argument("--legs", default="serve,sustained,batch,dropin,output,pond,config0,cpu",
                    help="comma list of the legs to run (default: all); without serve the line "
                         "carries no headline value (profiling runs of single legs)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the run, stream this many MB through the library's probe kernel "
                         "(reads and writes n doubles: a copy of known size, to scale rocprofv3 "
                         "FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 "
                                   "(sharding one grid or a batch over GPUs is not built)"}))
        sys.exit(2)

    pkg = load_package()
    cases = __import__("sw2d_efdc_amd.cases", fromlist=["config1"])

    # measured bandwidth ceiling beside the 8 TB/s vendor peak: a streaming
    # device copy of 1 GiB (read + write), 20 times
    copy_gbs = pkg.sw2d.stream_copy_gbs(1 << 30, 20)
    log(f"stream copy ceiling: {copy_gbs:.0f} GB/s")

    legs = set(args.legs.split(","))

    # ---- headline: the serving session ----
    if "serve" in legs:
        sv = serve_leg(pkg, cases, args.serve_seconds, args.serve_rate, args.serve_hours,
                       args.serve_precision, args.serve_active_cells, args.serve_tile_rows,
                       args.serve_err_sample, plain=args.serve_plain)
        log(f"serving: {sv['value']:.3e} cell-updates/s, p50 {sv['p50_time_to_hydrograph_s']:.2f} "
            f"s, p99 {sv['p99_time_to_hydrograph_s']:.2f} s")
    else:
        sv = {"value": None, "wall_s": 0.0, "busy_s": 0.0, "launches": 0, "requests": 0,
              "workload": "not run",
              "roofline": {"achieved": None, "frac": None, "bytes_per_cell_update": None,
                           "kernel_avg_us": None}}
    ms_per_step = 1e3 * sv["busy_s"] / max(sv["launches"], 1)
    rf = sv["roofline"]
    traffic_bpc = serve_traffic(args.serve_precision)
    cells_per_launch = (sv["value"] or 0.0) * sv["wall_s"] / max(sv["launches"], 1)
    issue = serve_issue(args.serve_precision)
    valu_issue = None
    if issue and sv["value"]:
        per, peak = issue
        valu_issue = {"wave_instructions_per_cell_update": per,
                      "peak_wave_instructions_per_s": peak,
                      "bound_cell_updates_per_s": peak / per,
                      "frac": sv["value"] / (peak / per),
                      "source": "rocprofv3 SQ_INSTS_VALU over a serving session; "
                                "profiles/serving_issue.json"}
    out = {
        "metric": METRIC,
        "value": sv["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32 state / f64 balance", "mixed": "f32 fluxes / f64 depth",
                  "fp64": "f64"}[args.serve_precision],
        "data": "synthetic (seeded DEMs and storms generated by sw2d-efdc_amd/cases.py)",
        "config": {
            "workload": f"serving: {sv['workload']}; the workload names 600 s of arrivals"
                        + (" (this window is reduced)" if args.serve_seconds < 600 else ""),
            "arrival_window_s": args.serve_seconds,
            "arrival_rate_per_s": args.serve_rate,
            "requests": sv["requests"],
            "precision": args.serve_precision,
            "step": "one batched launch: one iteration of every running request",
            "steps_warmup_apply_to": "the drop-in legs (timed iterations / untimed iterations "
                                     "after their wetted start); the serving session runs every "
                                     "request of its window to the end",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": rf["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": rf["frac"],
            "traffic": (traffic_bpc * cells_per_launch) if traffic_bpc else None,
            "traffic_bytes_per_cell_update": traffic_bpc,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over a serving session, "
                              "scaled by a copy of known size; profiles/serving_traffic.json",
            "bytes_per_cell_update": rf["bytes_per_cell_update"],
            "kernel_avg_us": rf["kernel_avg_us"],
            "stream_copy_GBs": copy_gbs,
            "frac_of_stream_copy": (rf["achieved"] / copy_gbs) if rf["achieved"] else None,
            # the bound the kernel actually meets: VALU issue (one wave64
            # instruction per SIMD per 4 cycles) at its measured instruction
            # count per cell-update
            "valu_issue": valu_issue,
        },
        "serving": sv,
    }
    if args.sustained_rate > 0 and "sustained" in legs:
        # the same request mix at a rate one GPU keeps up with: latency
        # without a standing queue
        out["serving_sustained"] = serve_leg(pkg, cases, args.sustained_seconds,
                                             args.sustained_rate, args.serve_hours,
                                             args.serve_precision, args.serve_active_cells,
                                             args.serve_tile_rows, 0, plain=False,
                                             label="sustained")
        log(f"sustained {args.sustained_rate}/s: p50 "
            f"{out['serving_sustained']['p50_time_to_hydrograph_s']:.2f} s")
    if args.batch_scenarios > 0 and "batch" in legs:
        out["batch"] = batch_leg(pkg, cases, args.batch_scenarios, args.batch_plain_sample)
    if args.dropin and "dropin" in legs:
        out["dropin"] = dropin_legs(pkg, cases, args.size, args.steps, args.warmup,
                                    args.wet_time, args.plain_steps)
    if args.output_every > 0 and "output" in legs:
        spec = cases.config1(rows=args.size, cols=args.size)
        out["output_path"] = output_leg(pkg, spec, args.output_iters, args.output_every)
        oc = out["output_path"]["output_cost_ms"]
        log(f"outputs: {oc['blocking']:.1f} ms each blocking, {oc['snapshot']:.1f} ms as snapshots")
    if args.pond_break > 0 and "pond" in legs:
        out["precision_study"] = pond_break_leg(pkg, cases, args.pond_break)
        log(f"pond break: fp32 max|dh| {out['precision_study']['fp32']['max_abs_dh_m']:.2e} m")
    c0_run = None
    if not args.no_cpu and args.config0 and "config0" in legs:
        out["config0"], c0_run = config0_leg(pkg, cases)
        log(f"reference 256^2 case: GPU {out['config0']['gpu_fp64_s']:.3f} s")
    if not args.no_cpu and "cpu" in legs:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads, {args.cpu_seconds} s")
        out["cpu_baseline"], c0_cpu = cpu_baseline(cases, args.cpu_seconds, threads,
                                                   args.serve_hours, c0_run)
        if c0_cpu is not None:
            out["config0"].update(c0_cpu)
            out["config0"]["speedup"] = c0_cpu["cpu_oracle_s"] / out["config0"]["gpu_fp64_s"]
            log(f"reference 256^2 case: CPU oracle {c0_cpu['cpu_oracle_s']:.2f} s, final h equal: "
                f"{c0_cpu['final_h_equal']}")
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
