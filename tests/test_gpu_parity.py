"""GPU parity: libsw2d (fused and plain kernels) against the CPU FP64 oracle.

FP64 bar: the state (h, um, vn) and the derived fields (uu, vv, uua, vva) are
equal value for value (IEEE ==, so +0 and -0 count as equal) to the oracle's
after every tested run, which also makes the wet/dry mask (h >= th) identical.
The oracle restates SW2D-EFDC.cu; see oracle/sw2d_oracle.c.
FP32 bar (precision form): max|dh| <= 1 mm against the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("h", "um", "vn", "uu", "vv", "uua", "vva")


def dem(cases, R, C, seed, dx=10.0, slope=0.01, amp=2.0, holes=False):
    r = np.arange(R, dtype=np.float64)[:, None]
    z = slope * dx * (R - 1 - r) + cases.fbm(R, C, 0.7, amp, seed)
    z = np.broadcast_to(z, (R, C)).copy()
    if holes:
        rng = np.random.default_rng(seed + 7)
        z[rng.random((R, C)) < 0.03] = -9999.0
        z[:, :2] = -9999.0
        z[R // 3, C // 4:C // 2] = -9999.0
    return z


def compare(gpu, ora, fields=FIELDS, tag=""):
    for f in fields:
        a, b = gpu.field(f), ora.field(f)
        if not np.array_equal(a, b):
            d = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
            i, j = d[0]
            raise AssertionError(
                f"{tag} field {f}: {len(d)} cells differ, first at ({i},{j}) gpu={a[i, j]!r} "
                f"oracle={b[i, j]!r} max|diff|={np.nanmax(np.abs(a - b)):.3e}")


def test_device_math_bitwise(pkg, po):
    sw = pkg.sw2d
    rng = np.random.default_rng(5)
    x = np.concatenate([10 ** rng.uniform(-5, 3, 20000), [1e-4, 1.0, 2.0, 0.5, 1e-300, 3e-320]])
    # pow: the table-driven range [2^-700, 2^500], its edges and the general path
    edges = [np.nextafter(b, d) for b in (2.0 ** -700, 2.0 ** 500) for d in (0, np.inf)]
    xp = np.concatenate([x, 10 ** rng.uniform(-320, 308, 20000), edges,
                         [2.0 ** -700, 2.0 ** 500, 0.0, -1.0, np.inf, np.nan]])
    gp = sw.device_math(0, xp)
    ep = po.pow1333(xp)
    assert np.array_equal(gp, ep, equal_nan=True), "pow1333 differs from the oracle"
    got = sw.device_math(0, x)
    exp = po.pow1333(x)
    assert np.array_equal(got, exp), "pow1333 differs from the oracle"
    t = rng.uniform(-700, 700, 20000)
    assert np.array_equal(sw.device_math(1, t), np.array([po.lib().orc_exp(v) for v in t]))
    assert np.array_equal(sw.device_math(2, x), np.array([po.lib().orc_log(v) for v in x]))
    # sqrt, division and fmax must be the IEEE operations (numpy rounds correctly)
    a = 10 ** rng.uniform(-300, 300, 20000)
    b = 10 ** rng.uniform(-300, 300, 20000) * rng.choice([-1, 1], 20000)
    assert np.array_equal(sw.device_math(3, a), np.sqrt(a))
    with np.errstate(over="ignore", under="ignore"):   # quotients past the range are part of the check
        q = a / b
    assert np.array_equal(sw.device_math(4, a, b), q)
    nanx = np.array([np.nan, 1.0, -2.0, np.nan])
    assert np.array_equal(sw.device_math(5, nanx, np.zeros(4)), np.array([0.0, 1.0, 0.0, 0.0]))


def test_cross_lane_moves(pkg):
    """The west/east neighbour moves of the fused kernel: element i receives
    element i-1 (resp. i+1) of its group of 64; the group's first (last) gets 0."""
    sw = pkg.sw2d
    x = np.arange(1, 257, dtype=np.float64) * 1.5 + 1e-300
    lane = np.arange(256) % 64
    west = np.where(lane == 0, 0.0, np.roll(x, 1))
    east = np.where(lane == 63, 0.0, np.roll(x, -1))
    assert np.array_equal(sw.device_math(6, x), west)
    assert np.array_equal(sw.device_math(7, x), east)


def pair(pkg, po, z, h0, dx, *, kernel="fused", tile_rows=0, dt=0.5, manning=0.03, **kw):
    g = pkg.Case(z, h0, dx, dt=dt, manning=manning, kernel=kernel, tile_rows=tile_rows, **kw)
    o = po.Oracle(z, h0, dx, dt=dt, manning=manning,
                  **{k: v for k, v in kw.items() if k in ("INT", "INF", "LWL", "EV_WL_min", "th",
                                                         "g", "nodata", "time0")})
    return g, o


@pytest.mark.parametrize("kernel,tile_rows", [("fused", 0), ("fused", 2), ("fused", 3),
                                              ("fused", 8), ("fused", 13),
                                              ("plain", 0)])
def test_reference_mode_parity(pkg, po, cases, kernel, tile_rows):
    """Rain series with interval boundaries where stream_flow divides 0/0 and
    x/0 (the reference's own NaN/inf forcing), dry start on a sloping fBm DEM,
    three 58-column strips, several tile heights."""
    R, C = 61, 130
    z = dem(cases, R, C, 11)
    h0 = np.zeros((R, C))
    g, o = pair(pkg, po, z, h0, 10.0, kernel=kernel, tile_rows=tile_rows)
    rain = np.array([60.0, 60.0, 0.0, 30.0, 0.0, 0.0])
    for c in (g, o):
        c.set_rain(rain, 100.0)
    for n in (1, 7, 150, 260):
        g.step(n)
        o.step(n)
        compare(g, o, tag=f"{kernel}/{tile_rows} after +{n}")
    assert g.time == o.time
    assert np.count_nonzero(o.field("h") > 1e-4) > 100  # the run did wet the domain


@pytest.mark.parametrize("kernel", ["fused", "plain"])
def test_holes_sources_evaporation(pkg, po, cases, kernel):
    """NODATA cells (inactive), a point source and a point sink, evaporation
    with a depth threshold, interception/infiltration/lake-loss fractions."""
    R, C = 47, 75
    dx = 10.0
    z = dem(cases, R, C, 3, holes=True)
    h0 = np.where(z > np.percentile(z[z > -9999], 30), 0.0, 0.05)
    h0[z == -9999.0] = 0.0
    kw = dict(INT=0.1, INF=0.05, LWL=0.2, EV_WL_min=0.02)
    g, o = pair(pkg, po, z, h0, dx, kernel=kernel, **kw)
    xs = np.array([30 * dx + 3.0, 12 * dx + 1.0])
    ys = np.array([(R - 20) * dx + 2.0, (R - 40) * dx - 4.0])
    Q = np.array([[0.5, -0.05], [1.5, -0.05], [0.0, -0.1], [2.0, -0.1]])
    cells = [po.source_index(R, C, 0.0, 0.0, dx, x, y) for x, y in zip(xs, ys)]
    rain = np.array([20.0, 5.0, 5.0, 0.0])
    evap = po.evaporation(0.2, np.array([20.0, 22.0, 25.0, 24.0]),
                          np.array([500.0, 800.0, 700.0, 100.0]), 200.0)
    assert np.allclose(pkg.sw2d.evaporation_series(0.2, [20.0, 22.0, 25.0, 24.0],
                                                   [500.0, 800.0, 700.0, 100.0], 200.0), evap,
                       rtol=0, atol=0)
    g.set_rain(rain, 200.0)
    o.set_rain(rain, 200.0)
    g.set_sources(xs, ys, Q, 150.0)
    o.set_sources(cells, Q, 150.0)
    g.set_evaporation(evap)
    o.set_evaporation(evap)
    for n in (3, 200, 250):
        g.step(n)
        o.step(n)
        compare(g, o, tag=f"{kernel} after +{n}")
    assert np.all(g.field("h")[z == -9999.0] == 0.0)


@pytest.mark.parametrize("kernel", ["fused", "plain"])
def test_extensions_parity(pkg, po, cases, kernel):
    """Adaptive CFL step, moving storm, Manning class map, outlet weir with
    hydrograph, lake-inflow records and the volume ledger."""
    spec = cases.config1(rows=96, cols=150, t_end=600.0)
    R, C = spec.z.shape
    g = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning, kernel=kernel,
                 adaptive=True, courant=spec.courant, tile_rows=16)
    o = po.Oracle(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
    st = dict(spec.storm)
    st.update(x0=0.0, y0=0.5 * R * spec.cellsize, vx=1.0, sigma=150.0, peak=200 / 3.6e6)
    for c in (g, o):
        c.set_storm(st["peak"], st["x0"], st["y0"], st["vx"], st["vy"], st["sigma"], st["t0"],
                    st["t1"])
        c.set_outlet(R - 1, int(np.argmin(spec.z[R - 1])))
        c.set_hydrograph(4096)
    g.set_manning_map(spec.nclass, spec.ntable)
    o.set_manning_map(spec.ntable[spec.nclass])
    o.set_adaptive(spec.courant, 2.0 * spec.dt)
    g.set_lake_bc(spec.lake_mask)
    rec = po.couple_bc(spec.lake_mask, np.zeros((R, C)), np.zeros((R, C)), np.zeros((R, C)),
                       spec.cellsize, spec.cellsize)
    cells = np.concatenate([rec[k]["cell"] for k in ("south", "north", "east", "west")])
    kinds = np.concatenate([np.full(len(rec[k]["cell"]), i) for i, k in
                            enumerate(("south", "north", "east", "west"))])
    o.set_bc(cells, kinds)
    assert g.nbc == len(cells) > 0
    for n in (5, 300, 400):
        g.step(n)
        o.step(n)
        compare(g, o, tag=f"{kernel} ext after +{n}")
        assert g.time == o.time
        assert g.clock()[2] == o.dt2_next
    hg, ho = g.hydrograph(), o.hydrograph()
    assert hg.shape == ho.shape and np.array_equal(hg, ho)
    assert ho[:, 1].max() > 0
    assert np.array_equal(g.lake_inflow(), o.bc_acc())
    bg, bo = g.balance(), o.ledger()
    for kg, ko in (("volume", "volume"), ("v_source", "v_src"), ("v_clip", "v_clip"),
                   ("v_outflow", "v_out"), ("q_peak", "q_peak"), ("t_peak", "t_peak")):
        assert bg[kg] == pytest.approx(bo[ko], rel=1e-12, abs=1e-9), kg
    assert abs(bg["error"]) <= 1e-9 * max(bg["volume"], 1.0)


def test_fp32_precision_form(pkg, po, cases):
    """FP32 state/fluxes with FP64 dt, volume and hydrograph against the FP64
    oracle: max|dh| <= 1 mm and volume error <= 1e-6 of the volume."""
    spec = cases.config0(rows=96, cols=96, outlet=True)
    g = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                 precision="fp32")
    o = po.Oracle(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning)
    for c in (g, o):
        c.set_rain(spec.rain * 4, spec.dtrain)
        c.set_outlet(*spec.outlet)
    g.step(1500)
    o.step(1500)
    dh = np.abs(g.field("h") - o.field("h"))
    assert dh.max() <= 1e-3, dh.max()
    b = g.balance()
    assert abs(b["error"]) <= 1e-6 * b["volume"]


def test_fp32_precision_form_with_holes(pkg, po, cases):
    """The FP32 form on a bed with NODATA cells (scattered, two edge columns
    and a row segment): inactive cells stay dry with no flow across their
    faces, and the depths follow the FP64 oracle within 1 mm.  The FP32 form
    reads per-cell bed steps whose NaN marks an inactive cell and whose value
    toward an inactive neighbour is never used."""
    R, C = 61, 83
    z = dem(cases, R, C, 11, holes=True)
    h0 = np.zeros((R, C))
    g = pkg.Case(z, h0, 10.0, dt=0.5, manning=0.03, precision="fp32")
    o = po.Oracle(z, h0, 10.0, dt=0.5, manning=0.03)
    rain = np.array([60.0, 30.0, 0.0])
    for c in (g, o):
        c.set_rain(rain, 300.0)
    g.step(900)
    o.step(900)
    hole = z == -9999.0
    h = g.field("h")
    assert np.all(h[hole] == 0.0)
    assert np.count_nonzero(o.field("h") > 1e-3) > 200   # the run wetted the domain
    dh = np.abs(h - o.field("h"))
    assert dh.max() <= 1e-3, dh.max()
    b = g.balance()
    assert abs(b["error"]) <= 1e-6 * b["volume"]


def test_autotune_keeps_results(pkg, po, cases):
    """The measured tile height changes the schedule, never the numbers."""
    R, C = 90, 140
    z = dem(cases, R, C, 21)
    h0 = np.zeros((R, C))
    g, o = pair(pkg, po, z, h0, 10.0)
    rain = np.array([80.0, 20.0, 0.0])
    g.set_rain(rain, 100.0)
    o.set_rain(rain, 100.0)
    H = g.autotune(3)
    assert 2 <= H <= 128
    g.step(120)
    o.step(120)
    compare(g, o, tag=f"autotuned H={H}")
    with pytest.raises(pkg.Sw2dError):
        g.autotune(0)            # at least one trial launch per candidate


def test_autotune_mid_run_keeps_results(pkg, po, cases):
    """Re-tuning the tile height after stepping has begun changes neither the
    state (value for value against the oracle) nor the ledger bookkeeping."""
    spec = cases.config0(rows=70, cols=131, seed=21, outlet=True)
    g, o = pair(pkg, po, spec.z, spec.h0, spec.cellsize, tile_rows=0, adaptive=True)
    for c in (g, o):
        c.set_rain(spec.rain, 600.0)
        c.set_outlet(*spec.outlet)
    o.set_adaptive(0.5, 1.0)
    g.set_hydrograph(400)
    o.set_hydrograph(400)
    # tilings with more and with fewer tiles than before, at odd and even
    # iterations (the live state is in either buffer)
    g.step(7)
    g.set_tile_rows(13)
    g.step(9)
    g.set_tile_rows(40)
    g.step(11)
    g.set_tile_rows(8)
    g.step(6)
    g.autotune(2)
    g.step(17)
    o.step(50)
    compare(g, o, tag="mid-run autotune")
    bg, bo = g.balance(), o.ledger()
    for kg, ko in (("volume", "volume"), ("v_source", "v_src"), ("v_clip", "v_clip"),
                   ("v_outflow", "v_out"), ("q_peak", "q_peak")):
        assert bg[kg] == pytest.approx(bo[ko], rel=1e-12, abs=1e-9), kg
    assert abs(bg["error"]) <= 1e-9 * bg["volume"]
    assert np.array_equal(g.hydrograph()[:, 1], o.hydrograph()[:, 1])


def test_periodic_retune_keeps_results(pkg, po, cases):
    """sw2d_set_retune: the tile height is re-timed inside step() every 25
    iterations (at odd and even iterations, inside and across graph groups);
    the state stays equal value for value to the oracle and the ledger and
    hydrograph to its."""
    spec = cases.config0(rows=90, cols=140, seed=23, outlet=True)
    g, o = pair(pkg, po, spec.z, spec.h0, spec.cellsize, tile_rows=0, adaptive=True)
    for c in (g, o):
        c.set_rain(spec.rain, 600.0)
        c.set_outlet(*spec.outlet)
    o.set_adaptive(0.5, 1.0)
    g.set_hydrograph(400)
    o.set_hydrograph(400)
    g.set_retune(25, 1)
    g.step(31)
    g.step(7)
    g.step(64)
    # a height below every round-filling candidate takes part in the next
    # retune (its scratch records are sized for it)
    g.set_tile_rows(5)
    g.step(30)
    g.set_retune(0)
    g.step(5)
    o.step(137)
    compare(g, o, tag="periodic retune")
    bg, bo = g.balance(), o.ledger()
    for kg, ko in (("volume", "volume"), ("v_source", "v_src"), ("v_clip", "v_clip"),
                   ("v_outflow", "v_out"), ("q_peak", "q_peak")):
        assert bg[kg] == pytest.approx(bo[ko], rel=1e-12, abs=1e-9), kg
    assert np.array_equal(g.hydrograph()[:, 1], o.hydrograph()[:, 1])


@pytest.mark.parametrize("feats", ["none", "map", "adaptive", "adaptive+map", "storm",
                                   "storm+adaptive+map", "sources"])
def test_kernel_instantiations_match_oracle(pkg, po, cases, feats):
    """The fused kernel is compiled once per feature set (storm, point
    sources, adaptive step, Manning map) and a launch runs the smallest
    instantiation covering its scenario.  Every combination is equal value for
    value to the oracle, and runs the instantiation the selection rule names."""
    R, C = 70, 125
    dx = 10.0
    z = dem(cases, R, C, 21)
    h0 = np.where(z < np.percentile(z, 25), 0.3, 0.0)
    on = set(feats.split("+"))
    g, o = pair(pkg, po, z, h0, dx, tile_rows=16, adaptive="adaptive" in on)
    if "adaptive" in on:
        o.set_adaptive(0.5, 2.0 * 0.5)
    rain = np.array([40.0, 10.0, 0.0])
    for c in (g, o):
        c.set_rain(rain, 300.0)
    if "map" in on:
        cls = (np.arange(R * C).reshape(R, C) // 7 % 3).astype(np.uint8)
        tab = np.array([0.02, 0.05, 0.09])
        g.set_manning_map(cls, tab)
        o.set_manning_map(tab[cls])
    if "storm" in on:
        for c in (g, o):
            c.set_storm(150 / 3.6e6, 0.0, 0.5 * R * dx, 2.0, 0.0, 200.0, 0.0, 600.0)
    if "sources" in on:
        xs, ys = np.array([40 * dx + 1.0]), np.array([(R - 30) * dx + 2.0])
        Q = np.array([[0.8], [1.2], [0.4]])
        g.set_sources(xs, ys, Q, 200.0)
        o.set_sources([po.source_index(R, C, 0.0, 0.0, dx, xs[0], ys[0])], Q, 200.0)
    want = ((pkg.sw2d.FEAT_STORM if "storm" in on else 0)
            | (pkg.sw2d.FEAT_SOURCES if "sources" in on else 0)
            | (pkg.sw2d.FEAT_ADAPTIVE if "adaptive" in on else 0)
            | (pkg.sw2d.FEAT_MAP if "map" in on else 0))
    variants = [0, 8, 12, 13, 15]
    f, v = g.kernel_features()
    assert f == want
    assert v == next(x for x in variants if f & ~x == 0)
    for n in (6, 180):
        g.step(n)
        o.step(n)
        compare(g, o, tag=f"{feats} after +{n}")
    assert g.time == o.time


def test_grid_smaller_than_2x2_is_rejected(pkg):
    for R, C in ((1, 1), (1, 70), (70, 1)):
        with pytest.raises(pkg.sw2d.Sw2dError, match="at least 2x2"):
            pkg.Case(np.zeros((R, C)), np.zeros((R, C)), 10.0, dt=0.5, manning=0.03)


@pytest.mark.parametrize("shape", [(2, 2), (2, 70), (70, 2), (2, 3), (3, 58), (4, 59), (6, 116),
                                   (7, 117), (17, 57), (33, 175)])
def test_grid_shapes_match_oracle(pkg, po, cases, shape):
    """Small and boundary shapes: two rows or two columns, widths at and
    around the 58-column strip (one lane short, exact, one over, two strips),
    heights around the tile height; rain onto a sloping bed with an outlet."""
    R, C = shape
    dx = 10.0
    z = dem(cases, R, C, 40 + R + C)
    h0 = np.where(z < np.median(z), 0.2, 0.0)
    for tile_rows in (8, 0):
        g, o = pair(pkg, po, z, h0, dx, tile_rows=tile_rows)
        for c in (g, o):
            c.set_rain(np.array([80.0, 20.0]), 120.0)
            c.set_outlet(R - 1, int(np.argmin(z[R - 1])))
        for n in (1, 5, 60):
            g.step(n)
            o.step(n)
            compare(g, o, tag=f"{shape} tile {tile_rows} after +{n}")
        assert g.time == o.time
        bg, bo = g.balance(), o.ledger()
        assert bg["volume"] == pytest.approx(bo["volume"], rel=1e-12, abs=1e-12)
        g.close()
