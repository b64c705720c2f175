"""Serving: continuous batching with HBM-footprint admission.

Each request's result (outlet hydrograph and volume ledger) must be the same
alone and in the mix: identical, value for value, to the same scenario run on
a stand-alone handle with the same tile rows, whatever ran beside it and
however long it queued.  The mix here arrives staggered while earlier
requests run, and the pool holds only part of it at once, so requests are
admitted as memory frees up.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 32
CAP = 4096   # hydrograph records kept per request


def request(cases, i):
    shapes = [(64, 90), (150, 120), (48, 300), (200, 200), (96, 96), (130, 70)]
    R, C = shapes[i % len(shapes)]
    rng = np.random.default_rng(500 + i)
    slope = rng.uniform(0.002, 0.03)
    r = np.arange(R, dtype=np.float64)[:, None]
    z = slope * 10.0 * (R - 1 - r) + cases.fbm(R, C, 0.7, rng.uniform(0.5, 5.0), 500 + i)
    z = np.broadcast_to(z, (R, C)).copy()
    outlet = (R - 1, int(np.argmin(z[R - 1])))
    rain = np.array([rng.uniform(10, 120) * 0.25, 0.0, 0.0, 0.0])   # mm per 15 min
    kw = dict(dt=0.5, manning=rng.uniform(0.02, 0.1), adaptive=bool(i % 3 == 2))
    t_end = 600.0 + 60.0 * i
    return z, outlet, rain, kw, t_end


def alone(pkg, z, outlet, rain, kw, t_end):
    c = pkg.Case(z, np.zeros_like(z), 10.0, precision="fp32", tile_rows=H, **kw)
    c.set_rain(rain, 900.0)
    c.set_outlet(*outlet)
    c.set_hydrograph(CAP)
    c.run_until(t_end)
    out = c.hydrograph(), c.balance()
    c.close()
    return out


def test_results_same_alone_and_in_the_mix(pkg, cases):
    n = 6
    reqs = [request(cases, i) for i in range(n)]
    fps = [pkg.footprint(z.shape[0], z.shape[1], precision="fp32", batch_tile_rows=H,
                         hydro_cap=CAP) for z, _, _, _, _ in reqs]
    # room for about half of the mix at once
    srv = pkg.Server(precision="fp32", pool_bytes=int(sum(sorted(fps)[-3:]) * 1.05), tile_rows=H,
                     slice_iters=12)
    ids = []
    for i, (z, outlet, rain, kw, t_end) in enumerate(reqs):
        ids.append(srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet,
                              hydro_cap=CAP, **kw))
        if i == 0:
            time.sleep(0.005)   # the first one starts alone; the others arrive while it runs
    srv.wait(timeout=300)
    st = srv.stats()
    assert st["completed"] == n and st["failed"] == 0 and st["queued"] == 0
    assert st["pool_peak"] <= st["pool_capacity"]
    res = [srv.poll(i) for i in ids]
    # overlap: some request was admitted before another one had finished
    assert any(a["t_admit"] < b["t_done"] and b["t_admit"] < a["t_done"]
               for a in res for b in res if a is not b)
    # the pool does not hold the whole mix: some request waited for a retirement
    assert any(a["t_admit"] >= b["t_done"] for a in res for b in res if a is not b)
    for i, (z, outlet, rain, kw, t_end) in enumerate(reqs):
        r = res[i]
        assert r["state"] == "done"
        hq, bal = alone(pkg, z, outlet, rain, kw, t_end)
        got = srv.hydrograph(ids[i])
        assert np.array_equal(got, hq), f"request {i} hydrograph differs from the stand-alone run"
        assert r["balance"] == bal, f"request {i} ledger differs"
        assert r["time"] <= t_end
        srv.release(ids[i])
    assert st["p99_latency"] >= st["p50_latency"] > 0
    srv.close()


def test_request_too_large_fails_cleanly(pkg):
    """A request larger than the whole pool fails; the server keeps serving."""
    big = np.zeros((600, 600)) + np.arange(600)[::-1, None] * 0.1   # ~10 MB of FP32 state
    srv = pkg.Server(precision="fp32", pool_bytes=2 << 20, tile_rows=H)
    rid = srv.submit(big, 10.0, t_end=60.0, dt=0.5, manning=0.03)
    srv.wait(rid, timeout=60)
    assert srv.poll(rid)["state"] == "failed"
    small = np.zeros((8, 8)) + np.arange(8)[::-1, None] * 0.1
    rid2 = srv.submit(small, 10.0, t_end=30.0, dt=0.5, manning=0.03)
    srv.wait(rid2, timeout=60)
    assert srv.poll(rid2)["state"] == "done"
    srv.close()


def test_active_cell_cap_orders_completion(pkg, cases):
    """With a cap on running cells the server admits the queue in arrival order
    as earlier requests finish; results stay those of the stand-alone runs."""
    reqs = [request(cases, i) for i in (0, 4, 5)]
    cells = [z.size for z, *_ in reqs]
    srv = pkg.Server(precision="fp32", pool_bytes=256 << 20, tile_rows=H, slice_iters=12,
                     max_active_cells=max(cells) + 1)
    ids = [srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet,
                      hydro_cap=CAP, **kw) for z, outlet, rain, kw, t_end in reqs]
    srv.wait(timeout=300)
    res = [srv.poll(i) for i in ids]
    assert all(r["state"] == "done" for r in res)
    # one at a time: each was admitted no earlier than the previous one finished
    for a, b in zip(res, res[1:]):
        assert b["t_admit"] >= a["t_done"]
    for i, (z, outlet, rain, kw, t_end) in enumerate(reqs):
        hq, bal = alone(pkg, z, outlet, rain, kw, t_end)
        assert np.array_equal(srv.hydrograph(ids[i]), hq)
    srv.close()


def test_manning_map_requests_same_alone_and_in_the_mix(pkg, cases):
    """Requests carrying a Manning class map run in a mix with requests that
    have none (and with adaptive ones), so the batch runs the kernel instance
    with the map; each hydrograph and ledger equals the stand-alone run's."""
    reqs = [request(cases, i) for i in (1, 2, 3)]
    maps = []
    for k, (z, *_rest) in enumerate(reqs):
        if k == 1:
            maps.append(None)
            continue
        R, C = z.shape
        cls = ((np.arange(R)[:, None] // 5 + np.arange(C)[None, :] // 7 + k) % 3).astype(np.uint8)
        maps.append((cls, np.array([0.025, 0.05, 0.09])))
    srv = pkg.Server(precision="fp32", pool_bytes=256 << 20, tile_rows=H, slice_iters=12)
    ids = [srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet,
                      hydro_cap=CAP, manning_map=mm, **kw)
           for (z, outlet, rain, kw, t_end), mm in zip(reqs, maps)]
    srv.wait(timeout=300)
    for i, ((z, outlet, rain, kw, t_end), mm) in enumerate(zip(reqs, maps)):
        r = srv.poll(ids[i])
        assert r["state"] == "done"
        c = pkg.Case(z, np.zeros_like(z), 10.0, precision="fp32", tile_rows=H, **kw)
        c.set_rain(rain, 900.0)
        c.set_outlet(*outlet)
        if mm is not None:
            c.set_manning_map(*mm)
        c.set_hydrograph(CAP)
        c.run_until(t_end)
        assert np.array_equal(srv.hydrograph(ids[i]), c.hydrograph()), f"request {i}"
        assert r["balance"] == c.balance(), f"request {i} ledger"
        c.close()
    srv.close()


def test_cancel_queued_and_running_requests(pkg, cases):
    """A queued request that is cancelled never runs; a running one leaves
    the batch after its slice and its cells and memory are released, so the
    request queued behind it is admitted and completes with the stand-alone
    result."""
    big = request(cases, 3)        # 200 x 200, the longest
    small = [request(cases, i) for i in (0, 4)]
    z, outlet, rain, kw, t_end = big
    # the cell cap admits the big request alone
    srv = pkg.Server(precision="fp32", pool_bytes=256 << 20, tile_rows=H, slice_iters=12,
                     max_active_cells=z.size)
    long_id = srv.submit(z, 10.0, t_end=20 * t_end, rain=rain, dtrain=900.0, outlet=outlet,
                         hydro_cap=CAP, **kw)
    ids = [srv.submit(z2, 10.0, t_end=t2, rain=r2, dtrain=900.0, outlet=o2, hydro_cap=CAP, **k2)
           for z2, o2, r2, k2, t2 in small]
    deadline = time.time() + 60
    while srv.poll(long_id)["state"] != "running" and time.time() < deadline:
        time.sleep(0.01)
    assert srv.poll(long_id)["state"] == "running"
    assert srv.poll(ids[0])["state"] == "queued"
    srv.cancel(ids[0])
    assert srv.poll(ids[0])["state"] == "cancelled"
    srv.cancel(long_id)
    srv.wait(timeout=300)
    st = srv.stats()
    assert st["cancelled"] == 2 and st["completed"] == 1 and st["failed"] == 0
    assert srv.poll(long_id)["state"] == "cancelled"
    assert srv.poll(ids[0])["iterations"] == 0
    r = srv.poll(ids[1])
    assert r["state"] == "done"
    z2, o2, r2, k2, t2 = small[1]
    hq, bal = alone(pkg, z2, o2, r2, k2, t2)
    assert np.array_equal(srv.hydrograph(ids[1]), hq)
    assert r["balance"] == bal
    with pytest.raises(pkg.sw2d.Sw2dError):
        srv.hydrograph(long_id)
    srv.release(long_id)
    srv.close()


def test_adaptive_tiling_keeps_results(pkg, cases):
    """tile_rows=0: the server retiles the running requests after admissions
    and retirements (grids near one round of wave slots at 2-row tiles, at
    most three running, so the height moves as the set changes while
    requests are mid-run).  Hydrographs equal the stand-alone runs value for
    value (the outlet is one cell of the state, which no tiling changes);
    the volume ledger, whose pending partial sums a retile folds, to
    rounding."""
    shapes = [(580, 580), (560, 600), (600, 540), (520, 640), (590, 570)]
    reqs = []
    for i, (R, C) in enumerate(shapes):
        rng = np.random.default_rng(900 + i)
        r = np.arange(R, dtype=np.float64)[:, None]
        z = 0.01 * 10.0 * (R - 1 - r) + cases.fbm(R, C, 0.7, rng.uniform(0.5, 3.0), 900 + i)
        z = np.broadcast_to(z, (R, C)).copy()
        outlet = (R - 1, int(np.argmin(z[R - 1])))
        rain = np.array([rng.uniform(20, 90) * 0.25, 0.0])
        kw = dict(dt=0.5, manning=rng.uniform(0.03, 0.08), adaptive=bool(i == 3))
        reqs.append((z, outlet, rain, kw, 150.0 + 60.0 * ((i * 3) % 5)))
    cap = int(3.2 * max(z.size for z, *_ in reqs))
    srv = pkg.Server(precision="fp32", pool_bytes=1 << 30, tile_rows=0, slice_iters=12,
                     max_active_cells=cap)
    ids = [srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet,
                      hydro_cap=CAP, **kw) for z, outlet, rain, kw, t_end in reqs]
    srv.wait(timeout=300)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["retiles"] >= 1 and 2 <= st["tile_rows"] <= 48
    for i, (z, outlet, rain, kw, t_end) in enumerate(reqs):
        r = srv.poll(ids[i])
        assert r["state"] == "done"
        hq, bal = alone(pkg, z, outlet, rain, kw, t_end)
        assert np.array_equal(srv.hydrograph(ids[i]), hq), f"request {i} hydrograph"
        for k in ("volume", "v_source", "v_clip", "v_outflow"):
            assert r["balance"][k] == pytest.approx(bal[k], rel=1e-12, abs=1e-9), (i, k)
    srv.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_server_with_adaptive_tiling_equals_alone(pkg, cases, precision):
    """The FP64 (parity) form and the mixed form (FP64 depths, FP32 fluxes)
    served with adaptive tiling: each hydrograph equals the stand-alone run's
    in the same form value for value."""
    reqs = [request(cases, i) for i in (0, 1, 2, 5)]
    srv = pkg.Server(precision=precision, pool_bytes=256 << 20, tile_rows=0, slice_iters=12,
                     max_active_cells=2 * max(z.size for z, *_ in reqs))
    ids = [srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet,
                      hydro_cap=CAP, **kw) for z, outlet, rain, kw, t_end in reqs]
    srv.wait(timeout=300)
    for i, (z, outlet, rain, kw, t_end) in enumerate(reqs):
        r = srv.poll(ids[i])
        assert r["state"] == "done"
        c = pkg.Case(z, np.zeros_like(z), 10.0, precision=precision, **kw)
        c.set_rain(rain, 900.0)
        c.set_outlet(*outlet)
        c.set_hydrograph(CAP)
        c.run_until(t_end)
        assert np.array_equal(srv.hydrograph(ids[i]), c.hydrograph()), f"request {i}"
        bal = c.balance()
        for k in ("volume", "v_source", "v_clip", "v_outflow"):
            assert r["balance"][k] == pytest.approx(bal[k], rel=1e-12, abs=1e-9), (i, k)
        c.close()
    srv.close()


def test_step_failure_fails_every_request_and_stops_admission(pkg, cases, monkeypatch):
    """A failed batch step (injected through SW2D_FAULT_SERVER_STEP: the
    third slice asks the batch for a negative iteration count, which
    sw2d_batch_step rejects) fails the running request and every queued one,
    with their finish times set, so a waiter returns; the server then refuses
    new requests instead of queueing work that would never run."""
    monkeypatch.setenv("SW2D_FAULT_SERVER_STEP", "2")
    reqs = [request(cases, i) for i in (0, 4, 5)]
    srv = pkg.Server(precision="fp32", pool_bytes=256 << 20, tile_rows=H, slice_iters=12,
                     max_active_cells=max(z.size for z, *_ in reqs) + 1)
    monkeypatch.delenv("SW2D_FAULT_SERVER_STEP")
    ids = [srv.submit(z, 10.0, t_end=20 * t_end, rain=rain, dtrain=900.0, outlet=outlet,
                      hydro_cap=CAP, **kw) for z, outlet, rain, kw, t_end in reqs]
    srv.wait(timeout=60)
    res = [srv.poll(i) for i in ids]
    assert [r["state"] for r in res] == ["failed"] * 3
    assert all(r["t_done"] > 0 and r["t_done"] >= r["t_submit"] for r in res)
    st = srv.stats()
    assert st["failed"] == 3 and st["running"] == 0 and st["queued"] == 0
    assert st["completed"] == 0
    z, outlet, rain, kw, t_end = reqs[0]
    with pytest.raises(pkg.sw2d.Sw2dError, match="batch step failed"):
        srv.submit(z, 10.0, t_end=t_end, rain=rain, dtrain=900.0, outlet=outlet, **kw)
    srv.close()
