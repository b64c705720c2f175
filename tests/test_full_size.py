"""The full-size workloads, through properties that do not depend on size.

The oracle finishes small grids in seconds, not the 2048^2 drop-in case or
the 4096^2 pond break.  At those sizes: the tiling does not change the
state (two tile heights give the same h, um, vn value for value, so every
strip and chunk seam of the real launch shape is exercised), the FP64 volume
ledger closes to round-off, and depths stay finite and non-negative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(pkg, spec, H, precision="fp64"):
    c = pkg.Case(spec.z, spec.h0, spec.cellsize, dt=spec.dt, manning=spec.manning,
                 precision=precision, adaptive=spec.adaptive, courant=spec.courant, tile_rows=H)
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


def check_properties(pkg, spec, heights, iters):
    states, times, bals = [], [], []
    for H in heights:
        c = make(pkg, spec, H)
        c.step(iters)
        h = c.field("h")
        assert np.isfinite(h).all() and (h >= 0).all(), f"H={H}"
        b = c.balance()
        assert abs(b["error"]) <= 1e-12 * b["volume"], (H, b)
        assert b["volume"] > 0
        states.append([c.field(n) for n in ("h", "um", "vn")])
        bals.append(b)
        times.append(c.time)
        c.close()
    ref = states[0]
    for H, st in zip(heights[1:], states[1:]):
        for name, a, b in zip(("h", "um", "vn"), ref, st):
            assert np.array_equal(a, b), f"{name} differs between tile heights {heights[0]} and {H}"
    assert len(set(times)) == 1
    return ref, bals[0]


def test_dropin_2048_tiling_invariance_and_ledger(pkg, cases):
    """2048^2 drop-in case (moving storm, Manning map, lake inflow rim,
    adaptive step), 150 iterations at tile heights 16 and 37."""
    spec = cases.config1(rows=2048, cols=2048)
    (h, um, vn), _ = check_properties(pkg, spec, (16, 37), 150)
    assert (h > spec.h0 + 1e-9).sum() > 1000   # the storm has wetted part of the plain


def test_pond_break_4096_tiling_invariance_and_ledger(pkg, cases):
    """4096^2 pond break (closed domain, adaptive step), 100 iterations at
    tile heights 24 and 48; no outflow, so the volume stays the initial one."""
    spec = cases.config2(rows=4096, cols=4096)
    (h, um, vn), b = check_properties(pkg, spec, (24, 48), 100)
    assert b["v_outflow"] == 0.0 and b["v_source"] == 0.0
    # the ledger's volume is the state's (clipping of negative depths at the
    # front is the only source of change, and the ledger carries it)
    assert b["volume"] == pytest.approx(h.sum() * spec.cellsize ** 2, rel=1e-12)
    assert b["volume0"] == pytest.approx(spec.h0.sum() * spec.cellsize ** 2, rel=1e-12)
    assert (h[spec.h0 == 0] > 0).any()           # the front has left the reservoir


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lake_at_rest_2048_exact(pkg, precision):
    """A flat free surface (eta = 10 m) over an uneven bed on a 1/8 m grid,
    no forcing, closed walls: every face flux stays exactly 0 and every depth
    exactly its initial value, at full size and in both forms (depths,
    surfaces and bed steps are exact binary fractions, so both the FP64
    surface test and the FP32 bed-step form see a zero gradient)."""
    R = C = 2048
    rng = np.random.default_rng(42)
    z = np.floor(rng.uniform(0.0, 8.0, (R, C)) * 8.0) / 8.0
    h0 = 10.0 - z
    c = pkg.Case(z, h0, 5.0, dt=0.5, manning=0.03, precision=precision)
    c.step(100)
    assert np.array_equal(c.field("h"), h0)
    assert not c.field("um").any() and not c.field("vn").any()
    b = c.balance()
    assert b["error"] == 0.0 or abs(b["error"]) <= 1e-15 * b["volume"]
    c.close()


def test_batch_of_large_uneven_grids_equals_stand_alone(pkg, cases):
    """Four large scenarios of uneven shapes (about 3.4 M cells together) in
    one batch, FP32 form, 60 iterations: each member's depths and fluxes equal
    its stand-alone run value for value (one launch's tile table then spans
    tens of thousands of blocks across the members)."""
    shapes = [(1024, 1024), (768, 1280), (512, 512), (2048, 256)]
    batch = pkg.Batch(precision="fp32", pool_bytes=2 << 30)
    members, alone = [], []
    for i, (R, C) in enumerate(shapes):
        r = np.arange(R, dtype=np.float64)[:, None]
        z = np.broadcast_to(0.01 * 10.0 * (R - 1 - r) + cases.fbm(R, C, 0.7, 1.5, 300 + i), (R, C)).copy()
        h0 = np.zeros_like(z)
        rain = np.array([40.0 + 10 * i, 20.0])
        a = pkg.Case(z, h0, 10.0, dt=0.5, manning=0.04, precision="fp32")
        a.set_rain(rain, 60.0)
        a.step(60)
        alone.append([a.field(n) for n in ("h", "um", "vn")])
        a.close()
        m = batch.case(z, h0, 10.0, dt=0.5, manning=0.04)
        m.set_rain(rain, 60.0)
        batch.attach(m, max_iters=60)
        members.append(m)
    while batch.step(6):
        pass
    for i, m in enumerate(members):
        for name, want in zip(("h", "um", "vn"), alone[i]):
            assert np.array_equal(m.field(name), want), (shapes[i], name)
    batch.close()
