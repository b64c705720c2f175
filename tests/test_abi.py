"""The C ABI library (CPU only: loads, exports, argument checks)."""
import ctypes as ct

import numpy as np


def test_library_exports_every_header_symbol(pkg):
    sw = pkg.sw2d
    L = sw.lib()
    names = sw.header_symbols()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.sw2d_version() == 1


def test_default_params(pkg):
    sw = pkg.sw2d
    p = sw.Params()
    sw.lib().sw2d_default_params(ct.byref(p))
    assert (p.g, p.th, p.precision, p.kernel, p.courant) == (9.80, 1.0e-4, 0, 0, 0.5)


def test_create_argument_errors(pkg):
    """Bad arguments are rejected before any device work, with a message."""
    sw = pkg.sw2d
    L = sw.lib()
    z = np.zeros(4)
    h = np.zeros(4)
    p = sw.Params()
    L.sw2d_default_params(ct.byref(p))
    out = ct.c_void_p()
    g = sw.GridDesc(1, 4, 0.0, 0.0, 10.0, -9999.0)          # one row
    assert L.sw2d_create(ct.byref(g), sw._dp(z), sw._dp(h), ct.byref(p), ct.byref(out)) == -1
    assert b"at least 2x2" in L.sw2d_last_error(None)
    g = sw.GridDesc(2, 2, 0.0, 0.0, 10.0, -9999.0)
    p.kernel, p.precision = 1, 1                              # plain kernels are FP64 only
    assert L.sw2d_create(ct.byref(g), sw._dp(z), sw._dp(h), ct.byref(p), ct.byref(out)) == -7
    p.kernel, p.precision = 0, 0
    p.dt = 0.0
    assert L.sw2d_create(ct.byref(g), sw._dp(z), sw._dp(h), ct.byref(p), ct.byref(out)) == -1
    assert L.sw2d_create(ct.byref(g), None, sw._dp(h), ct.byref(p), ct.byref(out)) == -1


def test_host_functions_reject_bad_arguments(pkg):
    L = pkg.sw2d.lib()
    assert L.sw2d_couple_bc(0, 0, 1.0, 1.0, None, None, None, None, 0, None, None, None, None,
                            None) == -1
    assert L.sw2d_lake_catchment_smooth(2, 2, None, None) == -1
    assert L.sw2d_device_math(0, 99, None, None, None, 1) == -1


def test_evaporation_series_matches_oracle(pkg, po):
    T = np.array([5.0, 12.5, 20.0, 31.0, -2.0])
    Rg = np.array([0.0, 120.0, 640.0, 900.0, 50.0])
    a = pkg.sw2d.evaporation_series(0.12, T, Rg, 3600.0)
    b = po.evaporation(0.12, T, Rg, 3600.0)
    assert np.array_equal(a, b)


def test_snapshot_argument_errors(pkg):
    """The asynchronous output calls reject a NULL handle before any device work."""
    L = pkg.sw2d.lib()
    t = ct.c_long(-1)
    assert L.sw2d_snapshot(None, 1, 0.0, ct.byref(t)) == -1
    p = ct.POINTER(ct.c_double)()
    assert L.sw2d_snapshot_field(None, 0, 0, ct.byref(p)) == -1
    assert L.sw2d_snapshot_release(None, 0) == -1


def test_bench_runs_on_one_gpu_only():
    """bench.py --gpus N with N > 1 exits with a message instead of measuring."""
    import json
    import pathlib
    import subprocess
    import sys
    root = pathlib.Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_create_rejects_grids_beyond_32bit_offsets(pkg):
    """Offsets in the kernels are 32-bit (byte offsets of doubles in the
    fused kernel): a grid of 2^29 or more cells (padded) is refused before
    any array is read or any device work."""
    sw = pkg.sw2d
    L = sw.lib()
    z = np.zeros(4)
    p = sw.Params()
    L.sw2d_default_params(ct.byref(p))
    out = ct.c_void_p()
    g = sw.GridDesc(46341, 46341, 0.0, 0.0, 10.0, -9999.0)   # 2.1e9 cells
    assert L.sw2d_create(ct.byref(g), sw._dp(z), sw._dp(z), ct.byref(p), ct.byref(out)) == -1
    assert b"too large" in L.sw2d_last_error(None)
    g = sw.GridDesc(16384, 32760, 0.0, 0.0, 10.0, -9999.0)   # pitch 32768: 2^29 padded cells
    assert L.sw2d_create(ct.byref(g), sw._dp(z), sw._dp(z), ct.byref(p), ct.byref(out)) == -1
