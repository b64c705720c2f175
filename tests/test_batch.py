"""Batches: scenarios of different sizes stepped by one launch per iteration.

The bar is identity with the same scenario stepped alone: every member keeps
the single-grid semantics, so its state (h, um, vn) after a batch run is
equal value for value to that of a stand-alone handle run for the same
iterations (same tile rows, so the ledger's summation order is the same too),
whatever else shares the launches and whenever it joined.  FP64 members are
also compared with the CPU FP64 oracle directly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("h", "um", "vn")


def dem(cases, R, C, seed, dx=10.0, slope=0.01, amp=2.0, holes=False):
    r = np.arange(R, dtype=np.float64)[:, None]
    z = slope * dx * (R - 1 - r) + cases.fbm(R, C, 0.7, amp, seed)
    z = np.broadcast_to(z, (R, C)).copy()
    if holes:
        rng = np.random.default_rng(seed + 3)
        z[rng.random((R, C)) < 0.02] = -9999.0
    return z


def scenario(cases, i):
    """A small mixed set: sizes, rain, outlet, Manning map, adaptive step."""
    shapes = [(40, 70), (96, 50), (130, 131), (23, 200), (64, 64)]
    R, C = shapes[i % len(shapes)]
    z = dem(cases, R, C, 100 + i, holes=(i == 2))
    h0 = np.zeros((R, C))
    if i == 4:
        h0[R // 4:R // 2, C // 4:C // 2] = 0.5
    kw = dict(adaptive=(i % 2 == 1), dt=0.5, manning=0.03 + 0.01 * i)
    rain = np.array([30.0 + 10 * i, 5.0, 0.0, 0.0])
    outlet = (R - 1, int(np.argmin(np.where(z[R - 1] > -9000, z[R - 1], np.inf))))
    return z, h0, kw, rain, outlet


def setup(case, cases, i, rain, outlet):
    case.set_rain(rain, 60.0)
    case.set_outlet(*outlet)
    case.set_hydrograph(4096)
    if i == 3:
        cls = (np.arange(case.R * case.C).reshape(case.R, case.C) % 3).astype(np.uint8)
        case.set_manning_map(cls, [0.02, 0.05, 0.08])


def assert_same(a, b, tag):
    for f in STATE:
        x, y = a.field(f), b.field(f)
        if not np.array_equal(x, y):
            d = np.argwhere(x != y)
            raise AssertionError(f"{tag} {f}: {len(d)} cells differ, first {tuple(d[0])} "
                                 f"{x[tuple(d[0])]!r} vs {y[tuple(d[0])]!r}")


@pytest.mark.parametrize("precision", ["fp32", "fp64", "mixed"])
def test_batch_equals_alone(pkg, cases, precision):
    H = 16
    iters = [37, 60, 25, 71, 48]
    join = [0, 0, 5, 13, 0]   # batch iterations before a member is attached (phases 5, 1)
    batch = pkg.Batch(precision=precision, pool_bytes=256 << 20, tile_rows=H)
    alone, members = [], []
    for i in range(5):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision=precision, tile_rows=H, **kw)
        setup(a, cases, i, rain, outlet)
        a.step(iters[i])
        alone.append(a)
        m = batch.case(z, h0, 10.0, **kw)
        setup(m, cases, i, rain, outlet)
        members.append(m)
    attached = [False] * 5
    k, chunk = 0, 0
    while True:
        for i in range(5):
            if not attached[i] and join[i] <= k:
                batch.attach(members[i], max_iters=iters[i])
                attached[i] = True
        n = (5, 8)[chunk % 2]   # ragged chunks: graph groups and single launches
        running = batch.step(n)
        k, chunk = k + n, chunk + 1
        for i in range(5):
            if attached[i]:
                assert batch.status(members[i])[1] <= iters[i]
        if all(attached) and running == 0:
            break
        assert k < 400
    for i in range(5):
        assert batch.status(members[i])[1] == iters[i]
        assert_same(members[i], alone[i], f"{precision} member {i}")
        assert members[i].time == alone[i].time
        ha, hm = alone[i].hydrograph(), members[i].hydrograph()
        assert np.array_equal(ha, hm), f"member {i} hydrograph"
        ba, bm = alone[i].balance(), members[i].balance()
        assert ba == bm, f"member {i} ledger"
    batch.close()


def test_batch_fp64_matches_oracle(pkg, po, cases):
    """FP64 members against the CPU FP64 oracle, value for value."""
    batch = pkg.Batch(precision="fp64", pool_bytes=128 << 20)
    pairs = []
    for i in (0, 2):
        z, h0, kw, rain, outlet = scenario(cases, i)
        kw.pop("adaptive")
        m = batch.case(z, h0, 10.0, **kw)
        m.set_rain(rain, 60.0)
        o = po.Oracle(z, h0, 10.0, dt=kw["dt"], manning=kw["manning"])
        o.set_rain(rain, 60.0)
        batch.attach(m, max_iters=90)
        pairs.append((m, o))
    while batch.step(12):
        pass
    for m, o in pairs:
        o.step(90)
        for f in ("h", "um", "vn", "uu", "vv", "uua", "vva"):
            assert np.array_equal(m.field(f), o.field(f)), f
    batch.close()


def test_batch_time_bound(pkg, cases):
    """t_end stops a member where run_until stops a stand-alone handle."""
    z, h0, kw, rain, outlet = scenario(cases, 1)   # adaptive
    a = pkg.Case(z, h0, 10.0, precision="fp32", tile_rows=32, **kw)
    a.set_rain(rain, 60.0)
    n_alone = a.run_until(95.0)
    batch = pkg.Batch(precision="fp32", pool_bytes=64 << 20, tile_rows=32)
    m = batch.case(z, h0, 10.0, **kw)
    m.set_rain(rain, 60.0)
    batch.attach(m, t_end=95.0)
    while batch.step(6):
        pass
    done, n, t = batch.status(m)
    assert done and n == n_alone and t == a.time and t <= 95.0
    assert_same(m, a, "t_end")
    batch.close()


def test_pool_admission(pkg, cases):
    """A member that does not fit the pool is refused with SW2D_E_NOMEM; it fits
    once another member is destroyed; the footprint bounds the pool use."""
    R, C = 300, 300
    fp = pkg.footprint(R, C, precision="fp32", batch_tile_rows=32, hydro_cap=100)
    batch = pkg.Batch(precision="fp32", pool_bytes=int(fp * 1.6), tile_rows=32)
    z = dem(cases, R, C, 5)
    h0 = np.zeros((R, C))
    a = batch.case(z, h0, 10.0, dt=0.5, manning=0.03)
    a.set_hydrograph(100)
    used, _, cap = batch.memory()
    assert used <= fp
    with pytest.raises(pkg.Sw2dError, match=r"\(-3\)"):
        batch.case(z, h0, 10.0, dt=0.5, manning=0.03)
    a.close()
    assert batch.memory()[0] == 0
    b = batch.case(z, h0, 10.0, dt=0.5, manning=0.03)
    batch.attach(b, max_iters=12)
    assert batch.step(12) == 0
    batch.close()


def test_batch_autotune_keeps_results(pkg, cases):
    batch = pkg.Batch(precision="fp32", pool_bytes=256 << 20)
    alone, members = [], []
    for i in range(3):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision="fp32", **kw)
        a.set_rain(rain, 60.0)
        a.step(30)
        alone.append(a)
        m = batch.case(z, h0, 10.0, **kw)
        m.set_rain(rain, 60.0)
        batch.attach(m, max_iters=30)
        members.append(m)
    H = batch.autotune(3)
    assert H > 0
    while batch.step(6):
        pass
    for i in range(3):
        assert_same(members[i], alone[i], f"autotuned member {i}")
    batch.close()


def test_batch_autotune_mid_run_keeps_results(pkg, cases):
    """Autotune at any point of a run: members at different iterations (one
    joined later, so its launch phase differs), trial launches at the live
    phase, then retiling with the pending partial sums folded.  The state
    equals the stand-alone runs and the volume ledgers agree to rounding."""
    batch = pkg.Batch(precision="fp32", pool_bytes=256 << 20)
    alone, members = [], []
    plan = [(0, 30), (0, 41), (7, 20), (0, 26)]   # (batch iterations before attach, iterations)
    for i, (join, n) in enumerate(plan):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision="fp32", **kw)
        setup(a, cases, i, rain, outlet)
        a.step(n)
        alone.append(a)
    for i, (join, n) in enumerate(plan):
        if join:
            continue
        z, h0, kw, rain, outlet = scenario(cases, i)
        m = batch.case(z, h0, 10.0, **kw)
        setup(m, cases, i, rain, outlet)
        batch.attach(m, max_iters=n)
        members.append((i, m))
    batch.step(7)
    z, h0, kw, rain, outlet = scenario(cases, 2)
    m = batch.case(z, h0, 10.0, **kw)
    setup(m, cases, 2, rain, outlet)
    batch.attach(m, max_iters=plan[2][1])
    members.append((2, m))
    batch.step(4)                 # launch phase 11 % 6 = 5
    H = batch.autotune(3)
    assert 2 <= H <= 128
    batch.step(3)
    H2 = batch.autotune(2)        # again, at another phase, after the retiling
    assert 2 <= H2 <= 128
    while batch.step(6):
        pass
    for i, m in members:
        assert_same(m, alone[i], f"member {i} autotuned mid-run")
        bm, ba = m.balance(), alone[i].balance()
        for k in ("volume", "v_source", "v_outflow"):
            assert bm[k] == pytest.approx(ba[k], rel=1e-12, abs=1e-9), (i, k)
        assert abs(bm["error"]) <= 1e-6 * max(bm["volume"], 1.0)
    batch.close()


def test_batch_autotune_with_a_member_of_fixed_small_tiles(pkg, cases):
    """A member created with tile_rows = 5 (below every autotune candidate)
    keeps its tiling through a batch autotune; the trials' scratch records
    cover its tile count.  Results equal the stand-alone runs."""
    batch = pkg.Batch(precision="fp32", pool_bytes=256 << 20)
    alone, members = [], []
    for i, rows in ((0, 5), (1, 0), (3, 0)):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision="fp32", **kw)
        setup(a, cases, i, rain, outlet)
        a.step(24)
        alone.append(a)
        m = batch.case(z, h0, 10.0, tile_rows=rows, **kw)
        setup(m, cases, i, rain, outlet)
        batch.attach(m, max_iters=24)
        members.append(m)
    batch.step(5)
    batch.autotune(2)
    while batch.step(6):
        pass
    for k, m in enumerate(members):
        assert_same(m, alone[k], f"member {k}")
    batch.close()


@pytest.mark.parametrize("order,waves", [(1, 0), (0, 2), (1, 1)])
def test_batch_schedule_keeps_results(pkg, cases, order, waves):
    """The launch schedule (block order, a cap on resident waves per SIMD)
    changes when and where tiles run, never the numbers: members equal
    their stand-alone runs value for value, set before the run and changed
    mid-run (graph replays captured afresh)."""
    batch = pkg.Batch(precision="fp32", pool_bytes=256 << 20, tile_rows=8)
    batch.set_schedule(order, waves)
    assert batch.schedule() == (order, waves)
    alone, members = [], []
    for i in range(4):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision="fp32", tile_rows=8, **kw)
        setup(a, cases, i, rain, outlet)
        a.step(40)
        alone.append(a)
        m = batch.case(z, h0, 10.0, **kw)
        setup(m, cases, i, rain, outlet)
        batch.attach(m, max_iters=40)
        members.append(m)
    batch.step(13)
    batch.set_schedule(1 - order, 0)
    while batch.step(6):
        pass
    for i in range(4):
        assert_same(members[i], alone[i], f"schedule ({order}, {waves}) member {i}")
        assert np.array_equal(members[i].hydrograph(), alone[i].hydrograph())
    batch.close()


def test_batch_autotune_searches_the_schedule(pkg, cases):
    """sw2d_batch_autotune times tile heights, then both block orders and
    caps below the register-limited occupancy at the chosen height; its
    report lists every trial and the pick is one of them; results are those
    of the stand-alone runs."""
    batch = pkg.Batch(precision="fp32", pool_bytes=256 << 20)
    alone, members = [], []
    for i in range(5):
        z, h0, kw, rain, outlet = scenario(cases, i)
        a = pkg.Case(z, h0, 10.0, precision="fp32", **kw)
        setup(a, cases, i, rain, outlet)
        alone.append(a)
        m = batch.case(z, h0, 10.0, **kw)
        setup(m, cases, i, rain, outlet)
        batch.attach(m, max_iters=36)
        members.append(m)
    batch.step(12)
    H = batch.autotune(2)
    rep = batch.tune_report()
    assert rep.shape[1] == 4 and len(rep) >= 4
    assert set(rep[:, 1].astype(int)) == {0, 1}                 # both orders tried
    assert (rep[:, 2] > 0).any()                                # some cap tried
    assert np.all(rep[:, 3] > 0)
    order, waves = batch.schedule()
    assert any(r[0] == H and r[1] == order and r[2] == waves for r in rep)
    while batch.step(6):
        pass
    for i in range(5):
        alone[i].step(36)
        # the state does not depend on the tiling or the schedule
        assert_same(members[i], alone[i], f"tuned member {i}")
    batch.close()


def test_eager_launches_equal_graph_replays(pkg, cases, monkeypatch):
    """SW2D_GRAPHS=0 at creation: every iteration is its own launch instead of
    a replay of the six-launch graph; a stand-alone case and a batch give the
    same state and ledger as with graph replays."""
    spec = cases.ensemble_scenario(5, side=160)

    def run(batch):
        if batch:
            b = pkg.Batch(precision="fp32", pool_bytes=128 << 20)
            m = b.case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
            m.set_rain(spec.rain, spec.dtrain)
            m.set_outlet(*spec.outlet)
            b.attach(m, t_end=spec.t_end)
            for _ in range(10):
                b.step(61)
            out = (m.field("h"), m.balance()["volume"], m.balance()["q_peak"])
            b.close()
            return out
        c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                     precision="fp64")
        c.set_rain(spec.rain, spec.dtrain)
        c.set_outlet(*spec.outlet)
        for _ in range(10):
            c.step(61)
        out = (c.field("h"), c.balance()["volume"], c.balance()["q_peak"])
        c.close()
        return out

    for batch in (False, True):
        monkeypatch.delenv("SW2D_GRAPHS", raising=False)
        g = run(batch)
        monkeypatch.setenv("SW2D_GRAPHS", "0")
        e = run(batch)
        monkeypatch.delenv("SW2D_GRAPHS")
        assert np.array_equal(g[0], e[0]) and g[1] == e[1] and g[2] == e[2]
        assert g[1] > 0
