"""The benchmark's workloads at their stated sizes, against the oracle (GPU).

- The reference's own CPU-runnable case (256 x 256, dx 10 m, 1 % slope + fBm
  2 m, n 0.03, 50 mm/h for 30 min then 30 min recession, fixed dt 0.5 s,
  outlet weir): the whole 3600-iteration run in FP64, state, hydrograph and
  ledger against the oracle's run of it.
- The serving stream (Poisson arrivals at 20/s, ensemble DEMs, 2 h block
  storms, fixed dt 0.5 s), with grid sides limited to those the oracle
  finishes in seconds (256 and 512): served together by an FP64 server, each
  request's hydrograph equals the oracle's run of that scenario value for
  value and its ledger to rounding.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def oracle_run(po, cases, spec, iters):
    o = po.Oracle(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                  threads=THREADS)
    cases.apply(o, spec)
    o.set_hydrograph(iters + 16)
    o.step(iters)
    return o


def ledger_matches(bal, led):
    for kg, ko in (("volume", "volume"), ("v_source", "v_src"), ("v_clip", "v_clip"),
                   ("v_outflow", "v_out"), ("q_peak", "q_peak"), ("t_peak", "t_peak")):
        assert bal[kg] == pytest.approx(led[ko], rel=1e-12, abs=1e-9), kg


def test_reference_case_whole_run_equals_oracle(pkg, po, cases):
    spec = cases.config0(outlet=True)
    iters = int(round(spec.t_end / (2 * spec.dt)))
    assert iters == 3600
    print(f"reference case {spec.z.shape}: {iters} iterations", flush=True)
    g = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                 precision="fp64")
    cases.apply(g, spec)
    g.set_hydrograph(iters + 16)
    g.step(iters)
    o = oracle_run(po, cases, spec, iters)
    assert g.time == o.time == spec.t_end
    for f in ("h", "um", "vn", "uu", "vv", "uua", "vva"):
        assert np.array_equal(g.field(f), o.field(f)), f
    hg, ho = g.hydrograph(), o.hydrograph()
    assert hg.shape == (iters, 2) and np.array_equal(hg, ho)
    assert ho[:, 1].max() > 0
    ledger_matches(g.balance(), o.ledger())
    g.close()


def test_serving_stream_equals_oracle(pkg, po, cases):
    arrivals, specs = cases.serve_requests(3, sides=(256, 512), probs=(0.6, 0.4))
    assert {s.z.shape[0] for s in specs} == {256, 512}
    iters = [int(round(s.t_end / (2 * s.dt))) for s in specs]
    cap = max(iters) + 16
    srv = pkg.Server(precision="fp64", pool_bytes=1 << 30, tile_rows=0, slice_iters=60)
    ids = []
    t0 = time.perf_counter()
    for a, s in zip(arrivals, specs):
        d = a - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        ids.append(srv.submit(s.z, s.cellsize, t_end=s.t_end, rain=s.rain, dtrain=s.dtrain,
                              outlet=s.outlet, hydro_cap=cap, dt=s.dt, manning=s.manning))
    srv.wait(timeout=600)
    st = srv.stats()
    assert st["completed"] == len(specs) and st["failed"] == 0
    for rid, s, k in zip(ids, specs, iters):
        r = srv.poll(rid)
        print(f"request {rid}: {s.z.shape} {k} iterations, q_peak "
              f"{r['balance']['q_peak']:.4f} m3/s", flush=True)
        assert r["state"] == "done" and r["iterations"] == k and r["time"] == s.t_end
        o = oracle_run(po, cases, s, k)
        hg = srv.hydrograph(rid)
        assert np.array_equal(hg, o.hydrograph()), f"request {rid} hydrograph"
        ledger_matches(r["balance"], o.ledger())
    srv.close()
