"""Pinning of the CPU FP64 oracle (oracle/sw2d_oracle.c) -- CPU only.

The reference holds no tests, fixtures or results for the SW2D path, so each
oracle function is pinned by what the mathematics fixes: library routines
(libm), closed forms, hand-computed branch values, invariants, and the
reference's own example input data (tests/golden/peri_lake.npz).
"""
import math
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
G = 9.80


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_pow1333_against_libm(po):
    rng = np.random.default_rng(1)
    x = np.concatenate([10 ** rng.uniform(-5, 3, 100000), [1e-4, 0.5, 1.0, 2.0, 1000.0]])
    got = po.pow1333(x)
    ref = np.power(x, 1.3333330)
    assert ulps(got, ref).max() <= 32
    assert po.pow1333([0.0])[0] == 0.0
    assert math.isnan(po.pow1333([-1.0])[0])


def pow_tables():
    """The three tables of sw2d_pow_tables.h (shared by the device pow and the
    oracle's), parsed from the header's hexadecimal constants."""
    import pathlib
    import re
    text = (pathlib.Path(__file__).resolve().parents[1] / "sw2d-efdc_amd" / "csrc" /
            "sw2d_pow_tables.h").read_text()
    out = {}
    for name in ("SW2D_POW_RECIP", "SW2D_POW_NLOG", "SW2D_POW_EXP2"):
        body = text.split(f"#define {name}")[1].split("\n\n")[0]
        out[name] = np.array([float.fromhex(h) for h in re.findall(r"-?0x[0-9a-fA-Fp.+-]+", body)])
        assert len(out[name]) == 64, name
    return out


def test_pow_tables():
    """RECIP[i] = RN(1 / (1 + (i + 1/2)/64)), NLOG[i] = RN(-ln RECIP[i]),
    EXP2[j] = RN(2^(j/64)): against IEEE division, libm's log and exp2 (within
    an ulp), the table's own identities, and monotonicity."""
    t = pow_tables()
    i = np.arange(64)
    recip, nlog, exp2 = t["SW2D_POW_RECIP"], t["SW2D_POW_NLOG"], t["SW2D_POW_EXP2"]
    assert np.array_equal(recip, 1.0 / (1.0 + (i + 0.5) / 64.0))
    # reciprocal of the bucket midpoint: r = m RECIP[i] - 1 stays below 2^-7
    lo, hi = 1.0 + i / 64.0, 1.0 + (i + 1) / 64.0
    assert np.all(np.abs(lo * recip - 1) < 2 ** -7) and np.all(np.abs(hi * recip - 1) < 2 ** -7)
    assert ulps(nlog, -np.log(recip)).max() <= 1
    assert np.all(np.diff(nlog) > 0) and nlog[0] > 0
    assert ulps(exp2, np.exp2(i / 64.0)).max() <= 1
    assert exp2[0] == 1.0 and np.all(np.diff(exp2) > 0) and exp2[-1] < 2.0
    # 2^(j/64) * 2^((64-j)/64) = 2, to rounding
    assert np.all(np.abs(exp2[1:] * exp2[:0:-1] - 2.0) <= 4 * np.spacing(2.0))


def test_pow1333_paths(po):
    """The table-driven range [2^-700, 2^500] and the general path beyond it
    agree where they meet, and the special values hold."""
    for b in (2.0 ** -700, 2.0 ** 500):
        xs = np.array([np.nextafter(b, 0), b, np.nextafter(b, np.inf)])
        got = po.pow1333(xs)
        ref = np.power(xs, 1.3333330)
        # |t| is large here (the rounding of t = 1.333 ln x dominates)
        assert (np.abs(got - ref) / np.abs(ref)).max() < 1e-13
    assert po.pow1333([1.0])[0] == 1.0
    assert po.pow1333([np.inf])[0] == np.inf
    assert math.isnan(po.pow1333([np.nan])[0])
    assert po.pow1333([1e-310])[0] == 0.0


def test_exp_log_against_libm(po):
    L = po.lib()
    rng = np.random.default_rng(2)
    t = rng.uniform(-700, 700, 50000)
    e = np.array([L.orc_exp(v) for v in t])
    assert ulps(e, np.exp(t)).max() <= 2
    x = 10 ** rng.uniform(-300, 300, 50000)
    lg = np.array([L.orc_log(v) for v in x])
    assert ulps(lg, np.log(x)).max() <= 2
    assert L.orc_exp(0.0) == 1.0 and L.orc_log(1.0) == 0.0


def test_forcing_interpolation_known_values(po):
    """stream_flow's rain term (SW2D-EFDC.cu:546-553), hand-evaluated."""
    rain, dtr = np.array([25.0, 0.0, 0.0, 0.0]), 1800.0
    # ramp for t <= 1: rain[0] * t
    assert po.rain_rate(rain, dtr, 0.0) == 0.0
    assert po.rain_rate(rain, dtr, 0.5) == pytest.approx(12.5 / 1.8e6, rel=1e-15)
    assert po.rain_rate(rain, dtr, 1.0) == pytest.approx(25.0 / 1.8e6, rel=1e-15)
    # t = 2: 25 + (0 - 25) / (1800 * 2) = 24.99305555...
    assert po.series(rain, dtr, 2.0) == pytest.approx(24.993055555555557, rel=1e-15)
    # exactly at t = 1800 the reference divides 0 by 0
    assert math.isnan(po.series(rain, dtr, 1800.0))
    # with unequal neighbours it divides a nonzero by 0: 10 -> 0 at t = 1800,
    # 0 -> 5 at t = 3600
    r2 = np.array([25.0, 10.0, 0.0, 5.0])
    assert po.series(r2, dtr, 1800.0) == -math.inf
    assert po.series(r2, dtr, 3600.0) == math.inf
    # between boundaries with equal neighbours the value is exact
    assert po.series(rain, dtr, 2000.0) == 0.0
    # past the end of the series the missing values read as 0
    assert po.series(np.array([3.0]), 10.0, 25.0) == 0.0


def test_source_index_example_case(po):
    """coord_source_sink.dat of the example case: (4921.04, 5215.8) m on the
    556 x 701, 10 m grid with origin (0, 0) -> qinx 492, qiny 701 - 522 = 179,
    flat index qiny * cols + qinx + 1."""
    idx = po.source_index(701, 556, 0.0, 0.0, 10.0, 4921.04, 5215.8)
    assert idx == 179 * 556 + 492 + 1


def test_evaporation_physical_values(po):
    """Energy-balance evaporation at 20 C: water density 998.2 kg/m3 (ITS-90)
    and latent heat 2.4535 MJ/kg (Henderson-Sellers 1984)."""
    E = po.evaporation(0.2, [20.0], [500.0], 3600.0)[0]
    expect = 400.0 / (2.4535e6 * 998.207) * 3600.0 * 1000.0
    assert E == pytest.approx(expect, rel=2e-4)
    assert po.evaporation(0.2, [20.0], [-50.0], 3600.0)[0] == 0.0


def two_cell(po, z, h, dt=0.5, n=0.03, **kw):
    """A 2 x 2 grid whose second column is NODATA: one row face between
    (0, 0) (north) and (1, 0) (south)."""
    zz = np.array([[z[0], -9999.0], [z[1], -9999.0]])
    hh = np.array([[h[0], 0.0], [h[1], 0.0]])
    return po.Oracle(zz, hh, 10.0, dt=dt, manning=n, **kw)


def test_dry_bed_weir_branch_1(po):
    """Surface of the south cell below the bed of the north cell: broad-crested
    weir 0.544 h sqrt(g h) from the north depth (SW2D-EFDC.cu:129-136)."""
    o = two_cell(po, (1.0, 0.0), (0.25, 0.5))
    o.step(1)
    assert o.field("um")[1, 0] == 0.544 * 0.25 * math.sqrt(G * 0.25)


def test_dry_bed_weir_branch_2(po):
    """Mirror case: north surface below the south bed, flux northwards."""
    o = two_cell(po, (0.0, 1.0), (0.5, 0.25))
    o.step(1)
    assert o.field("um")[1, 0] == -0.544 * 0.25 * math.sqrt(G * 0.25)


def test_dry_bed_weir_branch_3(po):
    """Depths straddling th: -sgn * 0.35 * hh3 * sqrt(2 g hh3) with
    hh3 = max surface - max bed = 0.2 - 0.1 (SW2D-EFDC.cu:148-158)."""
    o = two_cell(po, (0.0, 0.1), (0.2, 0.00005))
    o.step(1)
    assert o.field("um")[1, 0] == pytest.approx(0.35 * 0.1 * math.sqrt(2 * G * 0.1), rel=1e-15)


def test_treat_error_zeroes_outflow_of_drying_cell(po):
    """A shallow south cell (1.2e-4 m) drains through a weir face; with a long
    step its first-continuity depth falls below th, so treat_error zeroes the
    face and the second continuity leaves both depths unchanged."""
    o = two_cell(po, (0.0, 1.0), (0.0, 1.2e-4), dt=100.0)
    o.step(1)
    assert o.field("um")[1, 0] == 0.0
    h = o.field("h")
    assert h[1, 0] == 1.2e-4 and h[0, 0] == 0.0
    # deeper and shorter (dt2 = 20 s): no drying, the weir flux stands and
    # moves water north
    o = two_cell(po, (0.0, 1.0), (0.0, 0.01), dt=10.0)
    o.step(1)
    assert o.field("um")[1, 0] == -0.544 * 0.01 * math.sqrt(G * 0.01)
    assert o.field("h")[0, 0] > 0.0


def dyadic_lake(R=12, C=9, seed=3):
    rng = np.random.default_rng(seed)
    z = rng.integers(0, 40, size=(R, C)) / 8.0     # bed on a 1/8 m grid
    return z, 10.0 - z                             # flat surface at 10 m


def test_lake_at_rest(po):
    z, h = dyadic_lake()
    o = po.Oracle(z, h, 10.0, dt=0.5, manning=0.03)
    o.step(50)
    assert np.all(o.field("um") == 0) and np.all(o.field("vn") == 0)
    assert np.array_equal(o.field("h"), h)


def test_uniform_rain_on_flat_closed_box(po):
    """Flat closed box, constant rain series r mm per dtrain: every cell holds
    h_K = dt2 * rr * (K - 1/2) after K iterations (the second continuity runs
    at t + dt; t + dt = 0.5 is inside the start-up ramp rain[0] * t)."""
    R, C, r, dtr = 7, 11, 12.0, 600.0
    o = po.Oracle(np.zeros((R, C)), np.zeros((R, C)), 10.0, dt=0.5, manning=0.03)
    o.set_rain(np.full(10, r), dtr)
    K = 400
    o.step(K)
    rr = r / (dtr * 1000.0)
    h = o.field("h")
    assert np.all(h == h[0, 0])
    assert h[0, 0] == pytest.approx(1.0 * rr * (K - 0.5), rel=1e-12)
    led = o.ledger()
    assert led["volume"] == pytest.approx(led["v_src"], rel=1e-12)


def test_mass_conservation_closed_walls(po, cases):
    R, C = 40, 33
    z = cases.fbm(R, C, 0.7, 1.0, 4)
    h = np.zeros((R, C))
    h[10:20, 8:18] = 0.8
    o = po.Oracle(z, h, 5.0, dt=0.25, manning=0.03)
    v0 = h.sum()
    o.step(600)
    led = o.ledger()
    assert led["v_clip"] == 0.0
    assert o.field("h").sum() == pytest.approx(v0, rel=1e-13)
    assert np.abs(o.field("um")).max() > 0      # water did move


def test_kinematic_plane_steady_outflow(po):
    """Steady rain on a plane draining to a free-outflow outlet: the outlet
    discharge approaches rain * area (the ledger closes on the way)."""
    R, C, dx = 20, 3, 10.0
    r = np.arange(R, dtype=np.float64)[:, None]
    z = np.broadcast_to(0.05 * dx * (R - 1 - r), (R, C)).copy()
    z[:, 0] = 50.0
    z[:, 2] = 50.0                        # walls: a one-cell-wide channel
    o = po.Oracle(z, np.zeros((R, C)), dx, dt=0.25, manning=0.03)
    rain = np.full(40, 100.0)             # mm per 600 s
    o.set_rain(rain, 600.0)
    o.set_outlet(R - 1, 1)
    o.step(40000)
    led = o.ledger()
    area_rate = 100.0 / 600.0 / 1000.0 * dx * dx * R * C
    assert led["q_out"] == pytest.approx(area_rate, rel=0.02)
    err = led["volume"] - (led["v0"] + led["v_src"] + led["v_clip"] - led["v_out"])
    assert abs(err) <= 1e-9 * led["v_src"]


def test_adaptive_step_lake_at_rest(po):
    """At rest, dt2 = courant * dx / sqrt(g * hmax) on the deepest face, capped
    at dt2_max."""
    z, h = dyadic_lake()
    o = po.Oracle(z, h, 10.0, dt=5.0, manning=0.03)
    o.set_adaptive(0.5, 10.0)
    hm_max = max((0.5 * (h[1:, :] + h[:-1, :])).max(), (0.5 * (h[:, 1:] + h[:, :-1])).max())
    assert o.dt2_next == 0.5 * 10.0 / math.sqrt(G * hm_max)
    o.step(3)
    assert o.time == pytest.approx(3 * 0.5 * 10.0 / math.sqrt(G * hm_max), rel=1e-15)
    o2 = po.Oracle(z, h, 10.0, dt=0.1, manning=0.03)
    o2.set_adaptive(0.5, 0.2)
    assert o2.dt2_next == 0.2


def test_outlet_weir_discharge(po):
    """One wet cell on the south edge with the outlet: the first iteration's
    discharge is 0.5440 h sqrt(g h) * dy from the starting depth."""
    z = np.zeros((3, 3))
    h = np.zeros((3, 3))
    h[2, 1] = 0.3
    o = po.Oracle(z, h, 10.0, dt=0.5, manning=0.03)
    o.set_outlet(2, 1)
    o.set_hydrograph(10)
    o.step(1)
    q = 0.5440 * 0.3 * math.sqrt(G * 0.3) * 10.0
    hg = o.hydrograph()
    assert hg.shape == (1, 2) and hg[0, 0] == 1.0 and hg[0, 1] == q
    led = o.ledger()
    assert led["v_out"] == q * 1.0
    assert led["volume"] == pytest.approx(0.3 * 100.0 - q, rel=1e-12)


def numpy_couple_records(mask):
    """couple_BC_efdc's record sets written independently with numpy: for
    every catchment cell (mask != 1) the differing 4-neighbours it tests."""
    R, C = mask.shape
    recs = {"south": set(), "north": set(), "east": set(), "west": set()}
    cat = mask != 1
    ini = (R - 1) - np.arange(R)[:, None] * np.ones((1, C), int)   # scan row index per cell
    inj = np.ones((R, 1), int) * np.arange(C)[None, :]
    mid_r = (ini > 0) & (ini < R - 1)
    mid_c = (inj > 0) & (inj < C - 1)
    up = np.zeros_like(cat); up[1:, :] = mask[1:, :] != mask[:-1, :]        # p - C
    dn = np.zeros_like(cat); dn[:-1, :] = mask[:-1, :] != mask[1:, :]       # p + C
    lf = np.zeros_like(cat); lf[:, 1:] = mask[:, 1:] != mask[:, :-1]        # p - 1
    rt = np.zeros_like(cat); rt[:, :-1] = mask[:, :-1] != mask[:, 1:]       # p + 1
    tS = cat & up & ((mid_r) | (mid_c & (ini == 0)))
    tN = cat & dn & ((mid_r) | (mid_c & (ini == R - 1)))
    tE = cat & lf & ((mid_r & (mid_c | (inj == C - 1))) | (mid_c & ((ini == 0) | (ini == R - 1))))
    tW = cat & rt & ((mid_r & (mid_c | (inj == 0))) | (mid_c & ((ini == 0) | (ini == R - 1))))
    flat = np.arange(R * C).reshape(R, C)
    for name, t in (("south", tS), ("north", tN), ("east", tE), ("west", tW)):
        recs[name] = set(flat[t].tolist())
    return recs


def test_couple_bc_example_mask(po, pkg):
    """Records of couple_BC_efdc on the example case's lake mask (556 x 701,
    50638 lake cells): oracle, library and an independent numpy formulation
    agree; EFDC (I, J) follow the reference's per-edge offsets."""
    d = np.load(GOLDEN / "peri_lake.npz")
    mask = d["mask"]
    R, C = mask.shape
    assert (mask == 1).sum() == 50638
    rng = np.random.default_rng(0)
    h = rng.random((R, C))
    uua = rng.standard_normal((R, C))
    vva = rng.standard_normal((R, C))
    ro = po.couple_bc(mask, h, vva, uua, 10.0, 10.0)
    rl = pkg.couple_bc(mask, h, vva, uua, 10.0, 10.0)
    ind = numpy_couple_records(mask)
    for k in ("south", "north", "east", "west"):
        assert len(ro[k]["cell"]) > 0
        for f in ("cell", "I", "J", "q"):
            assert np.array_equal(ro[k][f], rl[k][f]), (k, f)
        assert set(ro[k]["cell"].tolist()) == ind[k], k
        cells = ro[k]["cell"]
        ini = (R - 1) - cells // C
        inj = cells % C
        off = {"south": (1, 2), "north": (1, 0), "east": (0, 1), "west": (2, 1)}[k]
        assert np.array_equal(ro[k]["I"], inj + off[0])
        assert np.array_equal(ro[k]["J"], ini + off[1])
        # scan order: south edge first, west to east
        assert np.all(np.diff(ini * C + inj) > 0)
    q = ro["north"]["q"]
    c = ro["north"]["cell"]
    u = uua.ravel()[c]
    assert np.array_equal(q, np.where(u > 0, u * h.ravel()[c] * 10.0, 0.0))


def test_lake_catchment_smooth_example_mask(po, pkg, cases):
    d = np.load(GOLDEN / "peri_lake.npz")
    mask = d["mask"]
    z = 20.0 + cases.fbm(*mask.shape, 0.7, 5.0, 9)
    z[0, :] = -9999.0
    a = po.lake_catchment_smooth(mask, z)
    b = pkg.lake_catchment_smooth(mask, z)
    assert np.array_equal(a, b)
    changed = a != z
    assert changed.sum() > 0 and np.all(mask[changed] != 1)


def test_derived_fields_of_linear_state(po):
    """hm_hn, uu1_vv1, uu_vv, uua_vva (SW2D-EFDC.cu:362-522) on states whose
    averages have closed forms: a field linear in (row, col) averaged over two
    or four face points equals its value at their centroid.  Dyadic
    coefficients keep every average exact, so a wrong neighbour index (a
    shifted centroid) changes the value.  Boundary rows/columns use the
    one-sided forms of the reference."""
    R, C = 9, 11
    r = np.arange(R, dtype=np.float64)[:, None] + np.zeros((1, C))
    j = np.arange(C, dtype=np.float64)[None, :] + np.zeros((R, 1))
    z = np.zeros((R, C))
    o = po.Oracle(z, np.ones((R, C)), 10.0, dt=0.5, manning=0.03)
    # depth linear in (r, j): face depths at the face centres
    h = 1.0 + 0.5 * r + 0.25 * j
    o.set_state(h, np.zeros((R, C)), np.zeros((R, C)))
    hm, hn = o.field("hm"), o.field("hn")
    assert np.array_equal(hm[1:], (1.0 + 0.5 * (r - 0.5) + 0.25 * j)[1:])
    assert np.array_equal(hm[0], h[0])                     # north edge: one-sided
    assert np.array_equal(hn[:, 1:], (1.0 + 0.5 * r + 0.25 * (j - 0.5))[:, 1:])
    assert np.array_equal(hn[:, 0], h[:, 0])               # west edge: one-sided
    # unit depth, fluxes linear in (r, j): velocities equal the fluxes
    a, b, c0 = 0.25, 0.125, 1.0
    um = a * r + b * j + c0
    vn = -0.5 * r + 0.375 * j - 2.0
    o.set_state(np.ones((R, C)), um, vn)
    uu1, vv1 = o.field("uu1"), o.field("vv1")
    assert np.array_equal(uu1, um) and np.array_equal(vv1, vn)
    uua, vva, uu, vv = o.field("uua"), o.field("vva"), o.field("uu"), o.field("vv")
    # uua: four row faces around the column face (r + 1/2, j - 1/2)
    assert np.array_equal(uua[:-1, 1:], (a * (r + 0.5) + b * (j - 0.5) + c0)[:-1, 1:])
    assert np.array_equal(uua[:-1, 0], (a * (r + 0.5) + c0)[:-1, 0])
    # vva: four column faces around the row face (r - 1/2, j + 1/2)
    lin_v = lambda rr, jj: -0.5 * rr + 0.375 * jj - 2.0  # noqa: E731
    assert np.array_equal(vva[1:, :-1], lin_v(r - 0.5, j + 0.5)[1:, :-1])
    assert np.array_equal(vva[0, :-1], lin_v(r, j + 0.5)[0, :-1])
    # cell centres
    assert np.array_equal(uu[:-1], (a * (r + 0.5) + b * j + c0)[:-1])
    assert np.array_equal(vv[:, :-1], lin_v(r, j + 0.5)[:, :-1])
    # the south row's uua and uu see the (zero) pad row below the grid
    assert np.array_equal(uu[-1], 0.5 * um[-1])
    # a dry cell (h < th) has zero centre velocity
    hd = np.ones((R, C))
    hd[4, 5] = 0.5e-4
    o.set_state(hd, um, vn)
    assert o.field("uu")[4, 5] == 0.0 and o.field("vv")[4, 5] == 0.0


# ---- B4 momentum update (SW2D-EFDC.cu:165-202 for um, 282-319 for vn) ----
# Each case sets a state whose one-iteration update has a closed form: the
# gravity term u13, the upwind convection terms u11/u12 (v11/v12) and the
# Manning friction ram, one at a time.  Dyadic inputs keep the convection
# cases exact, so a wrong neighbour, a central instead of an upwind
# difference or a flipped sign changes the result.

def momentum_step(po, h, um, vn, *, dx, dt=0.5, n=0.0):
    """One iteration on a flat bed from (h, um, vn); the new face fluxes."""
    R, C = h.shape
    o = po.Oracle(np.zeros((R, C)), h, dx, dt=dt, manning=n)
    o.set_state(h, um, vn)
    o.step(1)
    return o.field("um"), o.field("vn")


def grid_rc(R, C):
    r = np.arange(R, dtype=np.float64)[:, None] + np.zeros((1, C))
    j = np.arange(C, dtype=np.float64)[None, :] + np.zeros((R, 1))
    return r, j


@pytest.mark.parametrize("H,n,u,v", [(2.0, 0.05, 0.3, 0.0), (2.0, 0.05, 0.3, -0.2),
                                     (0.05, 0.08, 0.02, 0.015), (0.4, 0.03, -0.1, 0.25)])
def test_b4_friction_decay(po, H, n, u, v):
    """Uniform depth H on a flat bed with uniform face fluxes (u, v): the
    convection and gravity terms vanish and each face decays by the implicit
    Manning factor, um' = (1 - dt2 ram / 2) um / (1 + dt2 ram / 2) with
    ram = g n^2 |velocity| / H^1.3333330, |velocity| = hypot(u, v) / H (the
    cross velocity enters through vva for um and uua for vn)."""
    R, C = 8, 9
    h = np.full((R, C), H)
    um = np.full((R, C), u)
    um[0] = 0.0                         # the north edge face carries no flux
    vn = np.full((R, C), v)
    vn[:, 0] = 0.0                      # nor does the west edge face
    um1, vn1 = momentum_step(po, h, um, vn, dx=10.0, dt=0.5, n=n)
    dt2 = 1.0
    ram = G * n * n * (math.hypot(u, v) / H) / math.pow(H, 1.3333330)
    k = dt2 * ram * 0.5
    inner = np.s_[2:R - 1, 2:C - 1]     # faces whose stencils see only the uniform fluxes
    np.testing.assert_allclose(um1[inner], (1 - k) * u / (1 + k), rtol=1e-13, atol=0)
    np.testing.assert_allclose(vn1[inner], (1 - k) * v / (1 + k), rtol=1e-13, atol=0)
    # the decay is real: friction takes at least 0.01 % off each flux
    assert np.all(np.abs(um1[inner]) < np.abs(u) * (1 - 1e-4))


@pytest.mark.parametrize("s", [1.0, -1.0])
def test_b4_upwind_convection_along_flow(po, s):
    """n = 0, flat surface, um linear in the row, um = s (a + b r) with the
    flow all one way: the upwind flux difference is
    u11 = (dt2/dx) (2 beta m - s beta^2 / 2), beta = s b, m = um at the face.
    A central difference drops the beta^2 term, a downwind one flips it.
    The vn mirror v12 (vn linear in the column) has the same closed form."""
    R, C, dx = 10, 7, 8.0
    lam = 1.0 / dx                     # dt2 / dx with dt2 = 1
    r, j = grid_rc(R, C)
    a, b = 0.25, 0.125
    m = s * (a + b * r)
    um = m.copy()
    um[0] = 0.0
    um1, vn1 = momentum_step(po, np.ones((R, C)), um, np.zeros((R, C)), dx=dx)
    beta = s * b
    expect = m - lam * (2 * beta * m - s * beta * beta / 2)
    assert np.array_equal(um1[2:R - 1], expect[2:R - 1])
    assert np.all(vn1 == 0)
    # vn: linear in the column, the west edge face at 0
    w = s * (a + b * j)
    vn = w.copy()
    vn[:, 0] = 0.0
    um1, vn1 = momentum_step(po, np.ones((R, C)), np.zeros((R, C)), vn, dx=dx)
    expect = w - lam * (2 * beta * w - s * beta * beta / 2)
    assert np.array_equal(vn1[:, 2:C - 1], expect[:, 2:C - 1])
    assert np.all(um1 == 0)


@pytest.mark.parametrize("V", [0.5, -0.5])
def test_b4_upwind_cross_convection(po, V):
    """n = 0, flat surface.  um quadratic in the column, uniform down the rows,
    advected by a uniform cross velocity vn = V: u11 = 0 and
    u12 = (dt2/dy) V (um[j] - um[j-1]) for V > 0, (dt2/dy) V (um[j+1] - um[j])
    for V < 0.  The vn mirror: vn quadratic in the row advected by a uniform
    um = V gives v11 the same way (rows instead of columns)."""
    R, C, dx = 9, 10, 8.0
    lam = 1.0 / dx
    r, j = grid_rc(R, C)
    q = lambda x: 0.25 + 0.125 * x + x * x / 64.0    # noqa: E731  dyadic, positive
    um = q(j)
    um[0] = 0.0
    vn = np.full((R, C), V)
    vn[:, 0] = 0.0
    um1, vn1 = momentum_step(po, np.ones((R, C)), um, vn, dx=dx)
    jj = j[2:R - 1, 1:C - 1]
    diff = q(jj) - q(jj - 1) if V > 0 else q(jj + 1) - q(jj)
    assert np.array_equal(um1[2:R - 1, 1:C - 1], q(jj) - lam * V * diff)
    # the advecting flux itself sees no gradient: unchanged
    assert np.array_equal(vn1[2:R - 1, 2:C - 1], vn[2:R - 1, 2:C - 1])
    # vn mirror: v11 from a uniform row-direction velocity
    vn = q(r)
    vn[:, 0] = 0.0
    um = np.full((R, C), V)
    um[0] = 0.0
    um1, vn1 = momentum_step(po, np.ones((R, C)), um, vn, dx=dx)
    rr = r[1:R - 1, 2:C - 1]
    diff = q(rr) - q(rr - 1) if V > 0 else q(rr + 1) - q(rr)
    assert np.array_equal(vn1[1:R - 1, 2:C - 1], q(rr) - lam * V * diff)
    assert np.array_equal(um1[2:R - 1, 2:C - 1], um[2:R - 1, 2:C - 1])


def test_b4_gravity_term(po):
    """Fluid at rest under a sloping free surface over a flat bed: the first
    update is the pressure gradient alone, um' = -g hm (dt2/dx) dh with
    hm = the face depth (mean of the two cells) and dh = h(south) - h(north);
    likewise vn' = -g hn (dt2/dy) (h(east) - h(west)).  Flow runs down the
    surface: northwards (um < 0) when the surface rises southwards."""
    R, C, dx = 8, 9, 10.0
    r, j = grid_rc(R, C)
    s_r, s_c = 0.0625, -0.03125
    h = 1.0 + s_r * r + s_c * j
    um1, vn1 = momentum_step(po, h, np.zeros((R, C)), np.zeros((R, C)), dx=dx, n=0.03)
    lam = 1.0 / dx
    np.testing.assert_allclose(um1[1:], (-G * (h[1:] - 0.5 * s_r) * lam * s_r), rtol=1e-14)
    np.testing.assert_allclose(vn1[:, 1:], (-G * (h[:, 1:] - 0.5 * s_c) * lam * s_c), rtol=1e-14)
    assert np.all(um1[1:] < 0) and np.all(vn1[:, 1:] > 0)
    assert np.all(um1[0] == 0) and np.all(vn1[:, 0] == 0)


def test_manning_normal_depth_on_plane(po):
    """Steady rain i on a one-cell-wide plane of slope S draining to a free
    outlet (the second column is NODATA).  At steady state each row face
    carries the rain of the cells above it, um = i dx r, and gravity balances
    the reference's friction g n^2 |u| u / hm^(4/3) (ram), which is Manning's
    law um = hm^(5/3) sqrt(S_f) / n with the friction slope S_f.  Against the
    local surface slope S_f = -(eta_r - eta_(r-1)) / dx that balance holds to
    the convection term (< 1 %); against the bed slope S it is Manning's
    normal depth, within the few per cent the surface curvature adds.  A
    friction law with another depth exponent or n instead of n^2 misses by
    tens of per cent.  The plane is gentle enough (S dx = 1 cm, depths
    1.3-3.4 cm) that every face takes the momentum branch, not the dry-bed
    weirs (a surface below the next bed)."""
    R, dx, S, n, i_mm_h = 150, 2.0, 0.005, 0.03, 100.0
    r = np.arange(R, dtype=np.float64)
    z = np.stack([S * dx * (R - 1 - r), np.full(R, -9999.0)], axis=1)
    o = po.Oracle(z, np.zeros((R, 2)), dx, dt=0.5, manning=n)
    # one interval longer than the run: the series never reaches an interval
    # boundary (where stream_flow's interpolation divides 0 by 0)
    o.set_rain(np.full(3, i_mm_h * 100.0), 360000.0)
    o.set_outlet(R - 1, 0)
    o.step(10800)                         # 3 h; the wave crosses the plane in ~15 min
    i = i_mm_h / 1000.0 / 3600.0
    um, hm, h = o.field("um")[:, 0], o.field("hm")[:, 0], o.field("h")[:, 0]
    rows = np.arange(R // 5, R - 5)
    assert np.all(h[rows] > S * dx)       # no face is a weir
    np.testing.assert_allclose(um[rows], i * dx * rows, rtol=1e-9)
    eta = h + z[:, 0]
    s_f = -(eta[rows] - eta[rows - 1]) / dx
    np.testing.assert_allclose(um[rows], hm[rows] ** (5.0 / 3.0) * np.sqrt(s_f) / n, rtol=0.01)
    np.testing.assert_allclose(um[rows], hm[rows] ** (5.0 / 3.0) * math.sqrt(S) / n, rtol=0.03)
    led = o.ledger()
    assert led["q_out"] == pytest.approx(i * dx * dx * R, rel=1e-9)


def test_scheme_result_is_discontinuous_in_the_bed(po, cases):
    """The reference scheme switches discontinuously (a weir when a water
    surface falls below the neighbour's bed, momentum otherwise; the dry
    threshold), so on rough rained-on slopes its result is not a continuous
    function of its inputs.  Ensemble member 67 (157^2, dx 10 m, 1 h block
    storm): perturbing the bed by a relative 1e-15 moves no depth by more
    than 1e-12 m, but 1 um of seeded noise on the bed (far below survey
    accuracy) moves a pond's depth by several centimetres and the peak
    outflow by 0.1 % or more: water takes another path.  Whether a given
    perturbation flips a switch is a threshold event (some 10 nm draws do,
    others do not).  This is the yardstick for reduced-precision forms:
    rounding depths or fluxes to binary32 perturbs them by more than that."""
    s = cases.ensemble_scenario(67)
    iters = int(round(s.t_end / (2 * s.dt)))

    def run(z):
        o = po.Oracle(z, s.h0, s.cellsize, dt=s.dt, manning=s.manning, threads=4)
        o.set_rain(s.rain, s.dtrain)
        o.set_outlet(*s.outlet)
        o.set_hydrograph(iters + 8)
        o.step(iters)
        return o.field("h"), o.hydrograph()[:, 1].max()

    h0, q0 = run(s.z)
    h1, q1 = run(s.z * (1.0 + 1e-15 * np.random.default_rng(1).standard_normal(s.z.shape)))
    assert np.abs(h1 - h0).max() < 1e-12 and abs(q1 - q0) < 1e-12 * q0
    h2, q2 = run(s.z + 1e-6 * np.random.default_rng(3).standard_normal(s.z.shape))
    assert np.abs(h2 - h0).max() > 1e-2
    assert abs(q2 - q0) > 1e-4 * q0
