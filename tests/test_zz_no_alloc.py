"""sw2d_get_field_noalloc (GPU): the allocation-free variant of
sw2d_get_field gives the same values, value for value, on consecutive calls
of decreasing size, and makes no device allocation."""
import numpy as np
import pytest

FIELD_NAMES = ("h", "um", "vn", "uu", "vv", "uua", "vva", "z", "hm", "hn", "uu1", "vv1")


def wet_case(pkg, cases, R, C, precision, kernel, steps=30):
    r = np.arange(R, dtype=np.float64)[:, None]
    z = 0.2 * (R - 1 - r) + cases.fbm(R, C, 0.7, 2.0, 11)
    z[0, 0] = -9999.0       # an inactive cell: the z field maps it back to nodata
    c = pkg.Case(z, np.zeros((R, C)), 10.0, dt=0.5, manning=0.03, precision=precision,
                 kernel=kernel)
    c.set_rain(np.array([80.0, 20.0, 0.0]), 600.0)
    c.step(steps)
    c.sync()
    return c


@pytest.mark.gpu
def test_get_field_noalloc_equals_get_field(pkg, cases):
    # 2048 x 1100 = 2.25 M cells takes two chunks of the 2 Mi-value staging
    # area; the calls after it are smaller (a stale chunk would show)
    for R, C, prec, kern in ((2048, 1100, "fp64", "fused"), (90, 130, "fp32", "fused"),
                             (70, 50, "fp64", "plain")):
        print(f"get_field_noalloc: {R}x{C} {prec} {kern}", flush=True)
        c = wet_case(pkg, cases, R, C, prec, kern)
        before = c.device_bytes()
        got = {n: c.field_noalloc(n).copy() for n in FIELD_NAMES}
        assert c.device_bytes() == before          # no derive buffer was made
        assert got["h"].max() > 0                  # the run has water
        for n in FIELD_NAMES:
            assert np.array_equal(got[n], c.field(n)), (R, C, prec, kern, n)
        c.close()
