"""Asynchronous output copies (sw2d_snapshot*).

A snapshot holds the fields of the step it was taken at, whatever the handle
steps afterwards, and equals sw2d_get_field at that step value for value
(fused FP64, FP32 and the plain path).  Two snapshots can be held; a third
waits for a release and fails after its timeout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = ["h", "um", "vn", "uu", "vv", "uua", "vva", "z", "hm", "hn", "uu1", "vv1"]


def make(pkg, cases, **kw):
    R, C = 90, 131
    rng = np.random.default_rng(77)
    r = np.arange(R, dtype=np.float64)[:, None]
    z = np.broadcast_to(0.02 * 10.0 * (R - 1 - r) + cases.fbm(R, C, 0.7, 1.5, 77), (R, C)).copy()
    h0 = np.where(rng.random((R, C)) < 0.3, 0.05, 0.0)
    c = pkg.Case(z, h0, 10.0, dt=0.5, manning=0.04, **kw)
    c.set_rain(np.array([30.0, 5.0]), 60.0)
    c.set_outlet(R - 1, int(np.argmin(z[R - 1])))
    return c


@pytest.mark.parametrize("kw", [dict(precision="fp64"), dict(precision="fp32"),
                                dict(precision="mixed"), dict(precision="fp64", kernel="plain")],
                         ids=["fp64", "fp32", "mixed", "plain"])
def test_snapshot_equals_get_field_at_its_step(pkg, cases, kw):
    a, b = make(pkg, cases, **kw), make(pkg, cases, **kw)
    a.step(20)
    b.step(20)
    t1 = a.snapshot(ALL)
    want1 = {n: b.field(n) for n in ALL}
    a.step(15)                      # steps queued behind the snapshot's copies
    b.step(15)
    t2 = a.snapshot(["h", "uu"])
    want2 = {n: b.field(n) for n in ("h", "uu")}
    with pytest.raises(pkg.sw2d.Sw2dError):
        a.snapshot(["h"], timeout=0.0)     # both slots held
    a.step(7)
    for n in ALL:
        assert np.array_equal(a.snapshot_field(t1, n), want1[n]), n
    a.snapshot_release(t1)
    t3 = a.snapshot(["vv"])         # the released slot is reused
    b.step(7)
    assert np.array_equal(a.snapshot_field(t3, "vv"), b.field("vv"))
    for n in ("h", "uu"):
        assert np.array_equal(a.snapshot_field(t2, n), want2[n]), n
    with pytest.raises(pkg.sw2d.Sw2dError):
        a.snapshot_field(t2, "vn")  # not in that snapshot
    a.snapshot_release(t2)
    a.snapshot_release(t3)
    with pytest.raises(pkg.sw2d.Sw2dError):
        a.snapshot_release(t3)      # already released
    assert np.array_equal(a.field("h"), b.field("h"))
    a.close()
    b.close()


def test_snapshot_released_from_another_thread(pkg, cases):
    """A writer thread reads and releases snapshots while the caller keeps
    stepping and taking new ones (the output loop of the driver)."""
    import queue
    import threading

    a, b = make(pkg, cases, precision="fp32"), make(pkg, cases, precision="fp32")
    q = queue.Queue()
    got = {}

    def writer():
        while True:
            item = q.get()
            if item is None:
                return
            k, t = item
            got[k] = a.snapshot_field(t, "h")
            a.snapshot_release(t)

    th = threading.Thread(target=writer)
    th.start()
    want = {}
    for k in range(8):
        a.step(5)
        b.step(5)
        q.put((k, a.snapshot(["h"], timeout=60.0)))
        want[k] = b.field("h")
    q.put(None)
    th.join(timeout=120)
    assert not th.is_alive()
    for k in range(8):
        assert np.array_equal(got[k], want[k]), k
    a.close()
    b.close()


def test_snapshot_refused_for_batch_members(pkg, cases):
    batch = pkg.Batch(precision="fp32", pool_bytes=64 << 20)
    z = np.zeros((40, 40)) + np.arange(40)[::-1, None] * 0.1
    m = batch.case(z, np.zeros_like(z), 10.0, dt=0.5, manning=0.03)
    with pytest.raises(pkg.sw2d.Sw2dError):
        m.snapshot(["h"])
    batch.close()
