"""ctypes binding of libsw2d.so (the C ABI of include/sw2d.h).

This is the Python face of the library that the tests and the benchmark use.
It mirrors the C entry points one to one; there is no computation here and
no fallback: if libsw2d.so is missing or has no GPU, calls raise.
"""
import ctypes as ct
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsw2d.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sw2d.h")

FP64, FP32, MIXED = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}
# scenario features the fused kernel compiles out where unused (sw2d_kernel_features)
FEAT_STORM, FEAT_SOURCES, FEAT_ADAPTIVE, FEAT_MAP = 1, 2, 4, 8
FUSED, PLAIN = 0, 1
FIELDS = {"h": 0, "um": 1, "vn": 2, "uu": 3, "vv": 4, "uua": 5, "vva": 6, "z": 7, "hm": 8,
          "hn": 9, "uu1": 10, "vv1": 11}


class Sw2dError(RuntimeError):
    pass


class GridDesc(ct.Structure):
    _fields_ = [("rows", ct.c_int), ("cols", ct.c_int), ("xllcorner", ct.c_double),
                ("yllcorner", ct.c_double), ("cellsize", ct.c_double), ("nodata", ct.c_double)]


class Params(ct.Structure):
    _fields_ = [("g", ct.c_double), ("manning", ct.c_double), ("th", ct.c_double),
                ("dt", ct.c_double), ("time0", ct.c_double), ("INT", ct.c_double),
                ("INF", ct.c_double), ("LWL", ct.c_double), ("EV_WL_min", ct.c_double),
                ("precision", ct.c_int), ("kernel", ct.c_int), ("adaptive", ct.c_int),
                ("courant", ct.c_double), ("device", ct.c_int), ("tile_rows", ct.c_int)]


class Storm(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in ("peak", "x0", "y0", "vx", "vy", "sigma", "t0", "t1")]


class Balance(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in ("volume", "volume0", "v_source", "v_clip",
                                           "v_outflow", "q_out", "q_peak", "t_peak", "error")]


class Request(ct.Structure):
    _fields_ = [("grid", GridDesc), ("z", ct.POINTER(ct.c_double)), ("h0", ct.POINTER(ct.c_double)),
                ("params", Params), ("rain_mm", ct.POINTER(ct.c_double)), ("nrain", ct.c_int),
                ("dtrain", ct.c_double), ("outlet_row", ct.c_int), ("outlet_col", ct.c_int),
                ("manning_class", ct.POINTER(ct.c_ubyte)),
                ("manning_table", ct.POINTER(ct.c_double)), ("nclass", ct.c_int),
                ("t_end", ct.c_double), ("max_iters", ct.c_long), ("hydro_cap", ct.c_long)]


class Result(ct.Structure):
    _fields_ = [("state", ct.c_int), ("t_submit", ct.c_double), ("t_admit", ct.c_double),
                ("t_done", ct.c_double), ("iterations", ct.c_long), ("time", ct.c_double),
                ("hydro_n", ct.c_long), ("balance", Balance)]


class ServerStats(ct.Structure):
    _fields_ = [("submitted", ct.c_long), ("completed", ct.c_long), ("failed", ct.c_long),
                ("running", ct.c_long), ("queued", ct.c_long), ("cell_updates", ct.c_double),
                ("busy_seconds", ct.c_double), ("uptime", ct.c_double),
                ("p50_latency", ct.c_double), ("p99_latency", ct.c_double),
                ("max_latency", ct.c_double), ("p50_wait", ct.c_double), ("p99_wait", ct.c_double),
                ("pool_peak", ct.c_size_t), ("pool_capacity", ct.c_size_t),
                ("launches", ct.c_long), ("kernel_ms", ct.c_double), ("cancelled", ct.c_long),
                ("tile_rows", ct.c_long), ("retiles", ct.c_long)]


REQ_STATE = {0: "queued", 1: "running", 2: "done", -1: "failed", 3: "cancelled"}

_D = ct.c_double
_DP = ct.POINTER(ct.c_double)
_IP = ct.POINTER(ct.c_int)
_H = ct.c_void_p
_lib = None


def header_symbols():
    """Function names declared in include/sw2d.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sw2d_\w+)\s*\(", text, re.M)))


def lib():
    """Load libsw2d.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Sw2dError(f"{LIB_PATH} is missing: build the HIP library first "
                        "(python sw2d-efdc_amd/build.py)")
    L = ct.CDLL(LIB_PATH)
    sig = {
        "sw2d_default_params": (None, [ct.POINTER(Params)]),
        "sw2d_version": (ct.c_int, []),
        "sw2d_last_error": (ct.c_char_p, [_H]),
        "sw2d_create": (ct.c_int, [ct.POINTER(GridDesc), _DP, _DP, ct.POINTER(Params),
                                   ct.POINTER(_H)]),
        "sw2d_destroy": (ct.c_int, [_H]),
        "sw2d_set_rain_series": (ct.c_int, [_H, _DP, ct.c_int, _D]),
        "sw2d_set_sources": (ct.c_int, [_H, _DP, _DP, ct.c_int, _DP, ct.c_int, _D]),
        "sw2d_set_evaporation": (ct.c_int, [_H, _DP, ct.c_int]),
        "sw2d_evaporation_series": (ct.c_int, [_D, _DP, _DP, ct.c_int, _D, _DP]),
        "sw2d_set_manning_map": (ct.c_int, [_H, ct.POINTER(ct.c_ubyte), _DP, ct.c_int]),
        "sw2d_set_storm": (ct.c_int, [_H, ct.POINTER(Storm)]),
        "sw2d_set_outlet": (ct.c_int, [_H, ct.c_int, ct.c_int]),
        "sw2d_set_hydrograph": (ct.c_int, [_H, ct.c_long]),
        "sw2d_set_lake_bc": (ct.c_int, [_H, _IP]),
        "sw2d_autotune": (ct.c_int, [_H, ct.c_int, _IP]),
        "sw2d_set_tile_rows": (ct.c_int, [_H, ct.c_int]),
        "sw2d_step": (ct.c_int, [_H, ct.c_int]),
        "sw2d_run_until": (ct.c_int, [_H, _D, ct.c_long, ct.POINTER(ct.c_long)]),
        "sw2d_sync": (ct.c_int, [_H]),
        "sw2d_time": (ct.c_int, [_H, _DP, ct.POINTER(ct.c_long), _DP]),
        "sw2d_get_field": (ct.c_int, [_H, ct.c_int, _DP]),
        "sw2d_get_field_noalloc": (ct.c_int, [_H, ct.c_int, _DP]),
        "sw2d_batch_set_schedule": (ct.c_int, [_H, ct.c_int, ct.c_int]),
        "sw2d_batch_schedule": (ct.c_int, [_H, _IP, _IP]),
        "sw2d_batch_tune_report": (ct.c_int, [_H, _DP, ct.c_int, _IP]),
        "sw2d_snapshot": (ct.c_int, [_H, ct.c_uint, _D, ct.POINTER(ct.c_long)]),
        "sw2d_snapshot_field": (ct.c_int, [_H, ct.c_long, ct.c_int, ct.POINTER(_DP)]),
        "sw2d_snapshot_release": (ct.c_int, [_H, ct.c_long]),
        "sw2d_couple_bc": (ct.c_int, [ct.c_int, ct.c_int, _D, _D, _IP, _DP, _DP, _DP, ct.c_int,
                                      _IP, _IP, _IP, _DP, _IP]),
        "sw2d_lake_catchment_smooth": (ct.c_int, [ct.c_int, ct.c_int, _IP, _DP]),
        "sw2d_balance": (ct.c_int, [_H, ct.POINTER(Balance)]),
        "sw2d_hydrograph": (ct.c_int, [_H, _DP, ct.c_long, ct.POINTER(ct.c_long)]),
        "sw2d_lake_inflow": (ct.c_int, [_H, _DP, ct.c_int, _IP]),
        "sw2d_timing": (ct.c_int, [_H, ct.c_int]),
        "sw2d_kernel_time": (ct.c_int, [_H, _DP, ct.POINTER(ct.c_long)]),
        "sw2d_kernel_features": (ct.c_int, [_H, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
        "sw2d_set_retune": (ct.c_int, [_H, ct.c_long, ct.c_int]),
        "sw2d_stream_copy": (ct.c_int, [ct.c_int, ct.c_size_t, ct.c_int, _DP]),
        "sw2d_get_tile_rows": (ct.c_int, [_H, _IP]),
        "sw2d_device_bytes": (ct.c_int, [_H, ct.POINTER(ct.c_longlong)]),
        "sw2d_device_math": (ct.c_int, [ct.c_int, ct.c_int, _DP, _DP, _DP, ct.c_int]),
        "sw2d_batch_create": (ct.c_int, [ct.c_int, ct.c_int, ct.c_size_t, ct.c_int, ct.POINTER(_H)]),
        "sw2d_batch_destroy": (ct.c_int, [_H]),
        "sw2d_batch_last_error": (ct.c_char_p, [_H]),
        "sw2d_batch_footprint": (ct.c_size_t, [ct.POINTER(GridDesc), ct.POINTER(Params), ct.c_int,
                                               ct.c_long, ct.c_int]),
        "sw2d_batch_new": (ct.c_int, [_H, ct.POINTER(GridDesc), _DP, _DP, ct.POINTER(Params),
                                      ct.POINTER(_H)]),
        "sw2d_batch_attach": (ct.c_int, [_H, _H, ct.c_long, _D]),
        "sw2d_batch_detach": (ct.c_int, [_H, _H]),
        "sw2d_batch_step": (ct.c_int, [_H, ct.c_int, _IP]),
        "sw2d_batch_status": (ct.c_int, [_H, _H, _IP, ct.POINTER(ct.c_long), _DP]),
        "sw2d_batch_autotune": (ct.c_int, [_H, ct.c_int, _IP]),
        "sw2d_batch_memory": (ct.c_int, [_H, ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t),
                                         ct.POINTER(ct.c_size_t)]),
        "sw2d_batch_timing": (ct.c_int, [_H, ct.c_int]),
        "sw2d_batch_kernel_time": (ct.c_int, [_H, _DP, ct.POINTER(ct.c_long)]),
        "sw2d_server_create": (ct.c_int, [ct.c_int, ct.c_int, ct.c_size_t, ct.c_int, ct.c_int,
                                          ct.c_long, ct.POINTER(_H)]),
        "sw2d_server_submit": (ct.c_int, [_H, ct.POINTER(Request), ct.POINTER(ct.c_long)]),
        "sw2d_server_poll": (ct.c_int, [_H, ct.c_long, ct.POINTER(Result)]),
        "sw2d_server_hydrograph": (ct.c_int, [_H, ct.c_long, _DP, ct.c_long, ct.POINTER(ct.c_long)]),
        "sw2d_server_release": (ct.c_int, [_H, ct.c_long]),
        "sw2d_server_cancel": (ct.c_int, [_H, ct.c_long]),
        "sw2d_server_wait": (ct.c_int, [_H, ct.c_long, _D]),
        "sw2d_server_stats_get": (ct.c_int, [_H, ct.POINTER(ServerStats)]),
        "sw2d_server_last_error": (ct.c_char_p, [_H]),
        "sw2d_server_destroy": (ct.c_int, [_H]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def stream_copy_gbs(nbytes=1 << 30, reps=20, device=0):
    """Measured HBM ceiling: GB/s (read + write) of a streaming device copy."""
    g = ct.c_double(0)
    rc = lib().sw2d_stream_copy(device, nbytes, reps, ct.byref(g))
    if rc:
        raise Sw2dError(f"sw2d_stream_copy failed ({rc})")
    return g.value


def device_math(fn, x, y=None, device=0):
    """Evaluate the library's device math (0 pow1333, 1 exp, 2 log, 3 sqrt, 4 x/y, 5 fmax)."""
    x = _f64(np.ravel(x))
    yy = _f64(np.ravel(y)) if y is not None else None
    out = np.zeros_like(x)
    rc = lib().sw2d_device_math(device, fn, _dp(x), _dp(yy) if yy is not None else None,
                                _dp(out), len(x))
    if rc:
        raise Sw2dError(f"sw2d_device_math failed ({rc})")
    return out


def couple_bc(mask, h, vva, uua, dx, dy):
    """couple_BC_efdc records of a (rows, cols) mask: dict edge -> arrays."""
    R, C = mask.shape
    cap = R * C
    m = np.ascontiguousarray(mask, dtype=np.int32).ravel()
    cells = np.zeros(4 * cap, np.int32)
    I = np.zeros(4 * cap, np.int32)
    J = np.zeros(4 * cap, np.int32)
    q = np.zeros(4 * cap)
    counts = np.zeros(4, np.int32)
    rc = lib().sw2d_couple_bc(R, C, dx, dy, _ip(m), _dp(_f64(h).ravel()), _dp(_f64(vva).ravel()),
                              _dp(_f64(uua).ravel()), cap, _ip(cells), _ip(I), _ip(J), _dp(q),
                              _ip(counts))
    if rc:
        raise Sw2dError(f"sw2d_couple_bc failed ({rc})")
    out = {}
    for k, name in enumerate(("south", "north", "east", "west")):
        sl = slice(k * cap, k * cap + counts[k])
        out[name] = dict(cell=cells[sl].copy(), I=I[sl].copy(), J=J[sl].copy(), q=q[sl].copy())
    return out


def lake_catchment_smooth(mask, z):
    R, C = mask.shape
    m = np.ascontiguousarray(mask, dtype=np.int32).ravel()
    zz = _f64(z).ravel().copy()
    rc = lib().sw2d_lake_catchment_smooth(R, C, _ip(m), _dp(zz))
    if rc:
        raise Sw2dError(f"sw2d_lake_catchment_smooth failed ({rc})")
    return zz.reshape(R, C)


def evaporation_series(albedo, T, Rg, dtime):
    T = _f64(T)
    Rg = _f64(Rg)
    out = np.zeros(len(T))
    rc = lib().sw2d_evaporation_series(albedo, _dp(T), _dp(Rg), len(T), dtime, _dp(out))
    if rc:
        raise Sw2dError(f"sw2d_evaporation_series failed ({rc})")
    return out


def make_params(*, dt, manning, g=9.80, th=1.0e-4, INT=0.0, INF=0.0, LWL=0.0, EV_WL_min=0.0,
                time0=0.0, precision="fp64", kernel="fused", adaptive=False, courant=0.5,
                device=0, tile_rows=0):
    p = Params()
    lib().sw2d_default_params(ct.byref(p))
    p.g, p.manning, p.th, p.dt, p.time0 = g, manning, th, dt, time0
    p.INT, p.INF, p.LWL, p.EV_WL_min = INT, INF, LWL, EV_WL_min
    p.precision = PRECISIONS[precision]
    p.kernel = {"fused": FUSED, "plain": PLAIN}[kernel]
    p.adaptive = 1 if adaptive else 0
    p.courant = courant
    p.device = device
    p.tile_rows = tile_rows
    return p


def footprint(rows, cols, *, precision="fp32", tile_rows=0, hydro_cap=0, manning_map=False,
              batch_tile_rows=0):
    """Pool bytes a batch member of this shape takes (sw2d_batch_footprint)."""
    gd = GridDesc(rows, cols, 0.0, 0.0, 1.0, -9999.0)
    p = make_params(dt=0.5, manning=0.03, precision=precision, tile_rows=tile_rows)
    return int(lib().sw2d_batch_footprint(ct.byref(gd), ct.byref(p), batch_tile_rows, hydro_cap,
                                          1 if manning_map else 0))


class Case:
    """One SW2D case resident on the GPU (a sw2d_handle).  With batch=Batch(...)
    its arrays live in the batch's memory pool (sw2d_batch_new)."""

    def __init__(self, z, h0, cellsize, *, dt, manning, nodata=-9999.0, xll=0.0, yll=0.0,
                 batch=None, **kw):
        self._L = lib()
        z = _f64(z)
        h0 = _f64(h0)
        self.R, self.C = z.shape
        gd = GridDesc(self.R, self.C, xll, yll, cellsize, nodata)
        p = make_params(dt=dt, manning=manning, **kw)
        self._h = _H()
        self.batch = batch
        if batch is None:
            rc = self._L.sw2d_create(ct.byref(gd), _dp(z), _dp(h0), ct.byref(p), ct.byref(self._h))
            err = self._L.sw2d_last_error(None)
        else:
            rc = self._L.sw2d_batch_new(batch._b, ct.byref(gd), _dp(z), _dp(h0), ct.byref(p),
                                        ct.byref(self._h))
            err = self._L.sw2d_batch_last_error(batch._b)
        if rc:
            raise Sw2dError(f"sw2d_create failed ({rc}): {err.decode()}")
        if batch is not None:
            batch._cases.append(self)
        self.nbc = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.sw2d_destroy(self._h)
            self._h = None
            if self.batch is not None and self in self.batch._cases:
                self.batch._cases.remove(self)

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise Sw2dError(f"{what} failed ({rc}): {self._L.sw2d_last_error(self._h).decode()}")

    def set_rain(self, rain_mm, dtrain):
        r = _f64(rain_mm)
        self._chk(self._L.sw2d_set_rain_series(self._h, _dp(r), len(r), dtrain), "set_rain")

    def set_sources(self, x, y, Q, dtoq):
        x = _f64(x)
        y = _f64(y)
        Q = _f64(Q).reshape(-1, len(x)) if len(x) else np.zeros((0, 0))
        self._chk(self._L.sw2d_set_sources(self._h, _dp(x), _dp(y), len(x), _dp(Q), Q.shape[0],
                                           dtoq), "set_sources")

    def set_evaporation(self, evapo_mm):
        e = _f64(evapo_mm)
        self._chk(self._L.sw2d_set_evaporation(self._h, _dp(e), len(e)), "set_evaporation")

    def set_manning_map(self, cls, table):
        c = np.ascontiguousarray(cls, dtype=np.uint8)
        t = _f64(table)
        self._chk(self._L.sw2d_set_manning_map(
            self._h, c.ctypes.data_as(ct.POINTER(ct.c_ubyte)), _dp(t), len(t)), "set_manning_map")

    def set_storm(self, peak, x0, y0, vx, vy, sigma, t0, t1):
        s = Storm(peak, x0, y0, vx, vy, sigma, t0, t1)
        self._chk(self._L.sw2d_set_storm(self._h, ct.byref(s)), "set_storm")

    def set_outlet(self, r, c):
        self._chk(self._L.sw2d_set_outlet(self._h, r, c), "set_outlet")

    def set_hydrograph(self, cap):
        self._chk(self._L.sw2d_set_hydrograph(self._h, cap), "set_hydrograph")

    def set_lake_bc(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.int32)
        self._chk(self._L.sw2d_set_lake_bc(self._h, _ip(m)), "set_lake_bc")
        n = ct.c_int(0)
        self._chk(self._L.sw2d_lake_inflow(self._h, None, 0, ct.byref(n)), "lake_inflow")
        self.nbc = n.value

    def autotune(self, iters=20):
        """Measured choice of the fused kernel's tile height on the current state."""
        best = ct.c_int(0)
        self._chk(self._L.sw2d_autotune(self._h, iters, ct.byref(best)), "autotune")
        return best.value

    def set_tile_rows(self, rows):
        self._chk(self._L.sw2d_set_tile_rows(self._h, rows), "set_tile_rows")

    @property
    def tile_rows(self):
        """The fused kernel's current tile height (0 on the plain path)."""
        r = ct.c_int(0)
        self._chk(self._L.sw2d_get_tile_rows(self._h, ct.byref(r)), "get_tile_rows")
        return r.value

    def step(self, n=1):
        self._chk(self._L.sw2d_step(self._h, n), "step")

    def run_until(self, tmax, max_iters=1 << 62):
        done = ct.c_long(0)
        self._chk(self._L.sw2d_run_until(self._h, tmax, max_iters, ct.byref(done)), "run_until")
        return done.value

    def sync(self):
        self._chk(self._L.sw2d_sync(self._h), "sync")

    def clock(self):
        t = ct.c_double(0)
        m = ct.c_long(0)
        d = ct.c_double(0)
        self._chk(self._L.sw2d_time(self._h, ct.byref(t), ct.byref(m), ct.byref(d)), "time")
        return t.value, m.value, d.value

    @property
    def time(self):
        return self.clock()[0]

    def field(self, name, out=None):
        """sw2d_get_field; `out` (a C-contiguous float64 array of R*C values)
        is filled in place when given."""
        if out is None:
            out = np.zeros(self.R * self.C)
        if out.dtype != np.float64 or out.size != self.R * self.C or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of rows*cols values")
        self._chk(self._L.sw2d_get_field(self._h, FIELDS[name], _dp(out)), "get_field")
        return out.reshape(self.R, self.C)

    def field_noalloc(self, name, out=None):
        """sw2d_get_field_noalloc: the same values as field(), staged through
        the library's fixed per-device array (no device allocation)."""
        if out is None:
            out = np.zeros(self.R * self.C)
        if out.dtype != np.float64 or out.size != self.R * self.C or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of rows*cols values")
        self._chk(self._L.sw2d_get_field_noalloc(self._h, FIELDS[name], _dp(out)),
                  "get_field_noalloc")
        return out.reshape(self.R, self.C)

    def snapshot(self, names, timeout=60.0):
        """Queue an asynchronous copy of the named fields at the current step
        (sw2d_snapshot); returns its ticket."""
        mask = 0
        for n in names:
            mask |= 1 << FIELDS[n]
        t = ct.c_long(-1)
        self._chk(self._L.sw2d_snapshot(self._h, mask, timeout, ct.byref(t)), "snapshot")
        return t.value

    def snapshot_field(self, ticket, name, copy=True):
        """Wait for a snapshot's copy; returns the field, copied out of the
        pinned buffer, or with copy=False a read-only view of it that is valid
        until the ticket is released."""
        p = _DP()
        self._chk(self._L.sw2d_snapshot_field(self._h, ticket, FIELDS[name], ct.byref(p)),
                  "snapshot_field")
        a = np.ctypeslib.as_array(p, shape=(self.R * self.C,)).reshape(self.R, self.C)
        if copy:
            return a.copy()
        a.flags.writeable = False
        return a

    def snapshot_release(self, ticket):
        self._chk(self._L.sw2d_snapshot_release(self._h, ticket), "snapshot_release")

    def balance(self):
        b = Balance()
        self._chk(self._L.sw2d_balance(self._h, ct.byref(b)), "balance")
        return {n: getattr(b, n) for n, _ in Balance._fields_}

    def hydrograph(self, cap=1 << 20):
        out = np.zeros(2 * cap)
        n = ct.c_long(0)
        self._chk(self._L.sw2d_hydrograph(self._h, _dp(out), cap, ct.byref(n)), "hydrograph")
        m = min(n.value, cap)
        return out[: 2 * m].reshape(m, 2)

    def lake_inflow(self):
        out = np.zeros(max(self.nbc, 1))
        n = ct.c_int(0)
        self._chk(self._L.sw2d_lake_inflow(self._h, _dp(out), len(out), ct.byref(n)),
                  "lake_inflow")
        return out[: n.value]

    def timing(self, enable=True):
        self._chk(self._L.sw2d_timing(self._h, 1 if enable else 0), "timing")

    def kernel_time(self):
        ms = ct.c_double(0)
        n = ct.c_long(0)
        self._chk(self._L.sw2d_kernel_time(self._h, ct.byref(ms), ct.byref(n)), "kernel_time")
        return ms.value, n.value

    def set_retune(self, every, iters=3):
        """Re-time the tile height every `every` iterations inside step()
        (against the round-filling heights); 0 turns it off."""
        self._chk(self._L.sw2d_set_retune(self._h, int(every), int(iters)), "set_retune")

    def kernel_features(self):
        """(features, variant): SW2D_FEAT_* bits of this scenario and of the
        fused-kernel instantiation that runs it."""
        f = ct.c_int(0)
        v = ct.c_int(0)
        self._chk(self._L.sw2d_kernel_features(self._h, ct.byref(f), ct.byref(v)),
                  "kernel_features")
        return f.value, v.value

    def device_bytes(self):
        b = ct.c_longlong(0)
        self._chk(self._L.sw2d_device_bytes(self._h, ct.byref(b)), "device_bytes")
        return b.value


class Batch:
    """Scenarios stepped together by one launch per iteration (sw2d_batch_*)."""

    def __init__(self, precision="fp32", pool_bytes=1 << 30, tile_rows=0, device=0):
        self._L = lib()
        self._b = _H()
        self._cases = []
        rc = self._L.sw2d_batch_create(device, PRECISIONS[precision],
                                       pool_bytes, tile_rows, ct.byref(self._b))
        if rc:
            raise Sw2dError(f"sw2d_batch_create failed ({rc}): "
                            f"{self._L.sw2d_batch_last_error(None).decode()}")
        self.precision = precision

    def close(self):
        if getattr(self, "_b", None):
            for c in list(self._cases):
                c.close()
            self._L.sw2d_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise Sw2dError(f"{what} failed ({rc}): {self._L.sw2d_batch_last_error(self._b).decode()}")

    def case(self, z, h0, cellsize, **kw):
        kw.setdefault("precision", self.precision)
        return Case(z, h0, cellsize, batch=self, **kw)

    def attach(self, case, max_iters=-1, t_end=float("inf")):
        self._chk(self._L.sw2d_batch_attach(self._b, case._h, max_iters, t_end), "attach")

    def detach(self, case):
        self._chk(self._L.sw2d_batch_detach(self._b, case._h), "detach")

    def step(self, n=1):
        running = ct.c_int(0)
        self._chk(self._L.sw2d_batch_step(self._b, n, ct.byref(running)), "batch step")
        return running.value

    def status(self, case):
        d = ct.c_int(0)
        k = ct.c_long(0)
        t = ct.c_double(0)
        self._chk(self._L.sw2d_batch_status(self._b, case._h, ct.byref(d), ct.byref(k),
                                            ct.byref(t)), "status")
        return bool(d.value), k.value, t.value

    def autotune(self, iters=10):
        """Measured tile rows for the members, at any point of the run (trial
        launches of their next iteration; the ledgers are kept)."""
        best = ct.c_int(0)
        self._chk(self._L.sw2d_batch_autotune(self._b, iters, ct.byref(best)), "batch autotune")
        return best.value

    def set_schedule(self, order=0, waves_per_simd=0):
        """sw2d_batch_set_schedule: block order (0 runs, 1 interleaved) and a
        cap on resident waves per SIMD (0 = none)."""
        self._chk(self._L.sw2d_batch_set_schedule(self._b, order, waves_per_simd), "set_schedule")

    def schedule(self):
        o, w = ct.c_int(0), ct.c_int(0)
        self._chk(self._L.sw2d_batch_schedule(self._b, ct.byref(o), ct.byref(w)), "schedule")
        return o.value, w.value

    def tune_report(self):
        """Trials of the last autotune: rows (tile rows, order, waves cap, ms)."""
        n = ct.c_int(0)
        self._chk(self._L.sw2d_batch_tune_report(self._b, None, 0, ct.byref(n)), "tune_report")
        out = np.zeros(4 * max(n.value, 1))
        self._chk(self._L.sw2d_batch_tune_report(self._b, _dp(out), n.value, ct.byref(n)),
                  "tune_report")
        return out[: 4 * n.value].reshape(n.value, 4)

    def memory(self):
        a, b, c = ct.c_size_t(0), ct.c_size_t(0), ct.c_size_t(0)
        self._chk(self._L.sw2d_batch_memory(self._b, ct.byref(a), ct.byref(b), ct.byref(c)),
                  "memory")
        return a.value, b.value, c.value

    def timing(self, enable=True):
        self._chk(self._L.sw2d_batch_timing(self._b, 1 if enable else 0), "timing")

    def kernel_time(self):
        ms = ct.c_double(0)
        n = ct.c_long(0)
        self._chk(self._L.sw2d_batch_kernel_time(self._b, ct.byref(ms), ct.byref(n)),
                  "kernel_time")
        return ms.value, n.value


class Server:
    """Continuous-batching request loop on a device thread (sw2d_server_*).
    tile_rows=0 retiles the running requests after each admission or
    retirement to the height that fills whole rounds of wave slots."""

    def __init__(self, precision="fp32", pool_bytes=8 << 30, tile_rows=0, slice_iters=60,
                 max_active_cells=0, device=0):
        self._L = lib()
        self._s = _H()
        rc = self._L.sw2d_server_create(device, PRECISIONS[precision],
                                        pool_bytes, tile_rows, slice_iters, max_active_cells,
                                        ct.byref(self._s))
        if rc:
            raise Sw2dError(f"sw2d_server_create failed ({rc}): "
                            f"{self._L.sw2d_server_last_error(None).decode()}")

    def close(self):
        if getattr(self, "_s", None):
            self._L.sw2d_server_destroy(self._s)
            self._s = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise Sw2dError(f"{what} failed ({rc}): {self._L.sw2d_server_last_error(self._s).decode()}")

    def submit(self, z, cellsize, *, t_end, h0=None, rain=None, dtrain=1.0, outlet=None,
               manning_map=None, max_iters=-1, hydro_cap=0, nodata=-9999.0, **kw):
        z = _f64(z)
        R, C = z.shape
        q = Request()
        q.grid = GridDesc(R, C, 0.0, 0.0, cellsize, nodata)
        q.z = _dp(z)
        h0a = _f64(h0) if h0 is not None else None
        q.h0 = _dp(h0a) if h0a is not None else None
        q.params = make_params(**kw)
        ra = _f64(rain) if rain is not None else np.zeros(0)
        q.rain_mm = _dp(ra) if len(ra) else None
        q.nrain = len(ra)
        q.dtrain = dtrain
        q.outlet_row, q.outlet_col = outlet if outlet is not None else (-1, -1)
        if manning_map is not None:
            cls, tab = manning_map
            ca = np.ascontiguousarray(cls, dtype=np.uint8)
            ta = _f64(tab)
            q.manning_class = ca.ctypes.data_as(ct.POINTER(ct.c_ubyte))
            q.manning_table = _dp(ta)
            q.nclass = len(ta)
        q.t_end = t_end
        q.max_iters = max_iters
        q.hydro_cap = hydro_cap
        rid = ct.c_long(0)
        self._chk(self._L.sw2d_server_submit(self._s, ct.byref(q), ct.byref(rid)), "submit")
        return rid.value

    def poll(self, rid):
        r = Result()
        self._chk(self._L.sw2d_server_poll(self._s, rid, ct.byref(r)), "poll")
        out = {n: getattr(r, n) for n, _ in Result._fields_ if n != "balance"}
        out["state"] = REQ_STATE[r.state]
        out["balance"] = {n: getattr(r.balance, n) for n, _ in Balance._fields_}
        return out

    def hydrograph(self, rid):
        n = ct.c_long(0)
        self._chk(self._L.sw2d_server_hydrograph(self._s, rid, None, 0, ct.byref(n)), "hydrograph")
        out = np.zeros(2 * max(n.value, 1))
        self._chk(self._L.sw2d_server_hydrograph(self._s, rid, _dp(out), n.value, ct.byref(n)),
                  "hydrograph")
        return out[: 2 * n.value].reshape(n.value, 2)

    def release(self, rid):
        self._chk(self._L.sw2d_server_release(self._s, rid), "release")

    def cancel(self, rid):
        """Stop a queued or running request (its state becomes "cancelled")."""
        self._chk(self._L.sw2d_server_cancel(self._s, rid), "cancel")

    def wait(self, rid=-1, timeout=600.0):
        self._chk(self._L.sw2d_server_wait(self._s, rid, timeout), "wait")

    def stats(self):
        st = ServerStats()
        self._chk(self._L.sw2d_server_stats_get(self._s, ct.byref(st)), "stats")
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}
