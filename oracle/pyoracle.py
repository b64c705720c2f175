"""ctypes view of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline) of the HIP library, never as the product path.

The oracle restates the SW2D time step of SW2D-EFDC.cu (kernels 29-727, loop
4798-4877) in C; see oracle/sw2d_oracle.c for the per-function citations.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SW2D_ORACLE_LIB selects another build of the same source (the sanitizer
# build of `make -C oracle asan`)
LIB_PATH = os.environ.get("SW2D_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

F_H, F_UM, F_VN, F_UU, F_VV, F_UUA, F_VVA, F_Z, F_HM, F_HN, F_UU1, F_VV1, F_INF = range(13)
FIELDS = {"h": F_H, "um": F_UM, "vn": F_VN, "uu": F_UU, "vv": F_VV, "uua": F_UUA,
          "vva": F_VVA, "z": F_Z, "hm": F_HM, "hn": F_HN, "uu1": F_UU1, "vv1": F_VV1,
          "inf": F_INF}

_D = ct.c_double
_DP = ct.POINTER(ct.c_double)
_IP = ct.POINTER(ct.c_int)
_lib = None


def build():
    """Compile liboracle.so from oracle/sw2d_oracle.c (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        L.orc_create.restype = ct.c_void_p
        L.orc_create.argtypes = [ct.c_int, ct.c_int, _D, _D, _DP, _DP, _D, _D, _D, _D, _D, _D,
                                 _D, _D, _D]
        L.orc_destroy.argtypes = [ct.c_void_p]
        L.orc_set_threads.argtypes = [ct.c_void_p, ct.c_int]
        L.orc_set_rain.argtypes = [ct.c_void_p, _DP, ct.c_int, _D]
        L.orc_set_sources.argtypes = [ct.c_void_p, _IP, ct.c_int, _DP, ct.c_int, _D]
        L.orc_set_sources.restype = ct.c_int
        L.orc_set_evapo.argtypes = [ct.c_void_p, _DP, ct.c_int]
        L.orc_set_manning_map.argtypes = [ct.c_void_p, _DP]
        L.orc_set_storm.argtypes = [ct.c_void_p, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_set_outlet.argtypes = [ct.c_void_p, ct.c_int, ct.c_int]
        L.orc_set_outlet.restype = ct.c_int
        L.orc_set_hydrograph.argtypes = [ct.c_void_p, ct.c_long]
        L.orc_set_adaptive.argtypes = [ct.c_void_p, _D, _D]
        L.orc_set_bc.argtypes = [ct.c_void_p, _IP, _IP, ct.c_int]
        L.orc_step.argtypes = [ct.c_void_p, ct.c_int]
        L.orc_round_state_f32.argtypes = [ct.c_void_p]
        L.orc_set_state.argtypes = [ct.c_void_p, _DP, _DP, _DP]
        L.orc_set_state.restype = None
        L.orc_run_until.argtypes = [ct.c_void_p, _D, ct.c_long]
        L.orc_run_until.restype = ct.c_long
        L.orc_time.argtypes = [ct.c_void_p]
        L.orc_time.restype = _D
        L.orc_mstep.argtypes = [ct.c_void_p]
        L.orc_mstep.restype = ct.c_long
        L.orc_dt2_next.argtypes = [ct.c_void_p]
        L.orc_dt2_next.restype = _D
        L.orc_get_field.argtypes = [ct.c_void_p, ct.c_int, _DP]
        L.orc_get_field.restype = ct.c_int
        L.orc_ledger.argtypes = [ct.c_void_p, _DP]
        L.orc_hydrograph.argtypes = [ct.c_void_p, _DP, ct.c_long]
        L.orc_hydrograph.restype = ct.c_long
        L.orc_bc_acc.argtypes = [ct.c_void_p, _DP]
        for name in ("orc_pow1333", "orc_exp", "orc_log"):
            getattr(L, name).argtypes = [_D]
            getattr(L, name).restype = _D
        L.orc_series.argtypes = [_DP, ct.c_int, _D, _D]
        L.orc_series.restype = _D
        L.orc_rain_rate.argtypes = [_DP, ct.c_int, _D, _D]
        L.orc_rain_rate.restype = _D
        L.orc_source_index.argtypes = [ct.c_int, ct.c_int, _D, _D, _D, _D]
        L.orc_source_index.restype = ct.c_int
        L.orc_evaporation.argtypes = [_D, _DP, _DP, ct.c_int, _D, _DP]
        L.orc_bc_rate.argtypes = [ct.c_int, _D, _D, _D, _D, _D]
        L.orc_bc_rate.restype = _D
        L.orc_couple_bc.argtypes = [ct.c_int, ct.c_int, _D, _D, _IP, _DP, _DP, _DP, _IP, _IP,
                                    _IP, _DP, _IP, ct.c_int]
        L.orc_couple_bc.restype = ct.c_int
        L.orc_lake_catchment_smooth.argtypes = [ct.c_int, ct.c_int, _IP, _DP]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def pow1333(x):
    L = lib()
    return np.array([L.orc_pow1333(float(v)) for v in np.ravel(x)])


def series(v, dtser, t):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_series(_dp(v), len(v), float(dtser), float(t))


def rain_rate(v, dtrain, t):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_rain_rate(_dp(v), len(v), float(dtrain), float(t))


def source_index(rows, cols, xll, yll, res, x, y):
    return lib().orc_source_index(rows, cols, xll / res, yll / res, x / res, y / res)


def evaporation(albedo, T, Rg, dtime):
    T = np.ascontiguousarray(T, dtype=np.float64)
    Rg = np.ascontiguousarray(Rg, dtype=np.float64)
    out = np.zeros(len(T))
    lib().orc_evaporation(albedo, _dp(T), _dp(Rg), len(T), dtime, _dp(out))
    return out


def couple_bc(mask, h, vva, uua, dx, dy):
    """Records of couple_BC_efdc for a (rows, cols) mask: dict kind -> arrays."""
    R, C = mask.shape
    m = np.ascontiguousarray(mask, dtype=np.int32).ravel()
    hh = np.ascontiguousarray(h, dtype=np.float64).ravel()
    va = np.ascontiguousarray(vva, dtype=np.float64).ravel()
    ua = np.ascontiguousarray(uua, dtype=np.float64).ravel()
    cap = R * C
    cells = np.zeros(4 * cap, np.int32)
    I = np.zeros(4 * cap, np.int32)
    J = np.zeros(4 * cap, np.int32)
    q = np.zeros(4 * cap)
    counts = np.zeros(4, np.int32)
    rc = lib().orc_couple_bc(C, R, dx, dy, _ip(m), _dp(hh), _dp(va), _dp(ua), _ip(cells),
                             _ip(I), _ip(J), _dp(q), _ip(counts), cap)
    assert rc == 0
    out = {}
    for k, name in enumerate(("south", "north", "east", "west")):
        n = counts[k]
        sl = slice(k * cap, k * cap + n)
        out[name] = dict(cell=cells[sl].copy(), I=I[sl].copy(), J=J[sl].copy(), q=q[sl].copy())
    return out


def lake_catchment_smooth(mask, z):
    R, C = mask.shape
    m = np.ascontiguousarray(mask, dtype=np.int32).ravel()
    zz = np.ascontiguousarray(z, dtype=np.float64).ravel().copy()
    lib().orc_lake_catchment_smooth(C, R, _ip(m), _dp(zz))
    return zz.reshape(R, C)


class Oracle:
    """One SW2D case on the CPU (FP64).  Fields come back as (rows, cols)."""

    def __init__(self, z, h0, cellsize, *, dt, manning, nodata=-9999.0, g=9.80, th=1.0e-4,
                 INT=0.0, INF=0.0, LWL=0.0, EV_WL_min=0.0, time0=0.0, threads=1):
        z = np.ascontiguousarray(z, dtype=np.float64)
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        self.R, self.C = z.shape
        self._L = lib()
        self._s = self._L.orc_create(self.R, self.C, cellsize, nodata, _dp(z), _dp(h0), g,
                                     manning, th, dt, INT, INF, LWL, EV_WL_min, time0)
        self._L.orc_set_threads(self._s, threads)
        self.nbc = 0

    def __del__(self):
        if getattr(self, "_s", None):
            self._L.orc_destroy(self._s)
            self._s = None

    def set_rain(self, rain, dtrain):
        r = np.ascontiguousarray(rain, dtype=np.float64)
        self._L.orc_set_rain(self._s, _dp(r), len(r), dtrain)

    def set_sources(self, cells, Q, dtoq):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1, len(cells))
        rc = self._L.orc_set_sources(self._s, _ip(cells), len(cells), _dp(Q), Q.shape[0], dtoq)
        assert rc == 0

    def set_evaporation(self, evapo):
        e = np.ascontiguousarray(evapo, dtype=np.float64)
        self._L.orc_set_evapo(self._s, _dp(e), len(e))

    def set_manning_map(self, nmap):
        n = np.ascontiguousarray(nmap, dtype=np.float64)
        self._L.orc_set_manning_map(self._s, _dp(n))

    def set_storm(self, peak, x0, y0, vx, vy, sigma, t0, t1):
        self._L.orc_set_storm(self._s, peak, x0, y0, vx, vy, sigma, t0, t1)

    def set_outlet(self, r, c):
        assert self._L.orc_set_outlet(self._s, r, c) == 0

    def set_hydrograph(self, cap):
        self._L.orc_set_hydrograph(self._s, cap)

    def set_adaptive(self, courant, dt2_max):
        self._L.orc_set_adaptive(self._s, courant, dt2_max)

    def set_bc(self, cells, kinds):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        self._L.orc_set_bc(self._s, _ip(cells), _ip(kinds), len(cells))
        self.nbc = len(cells)

    def step(self, n=1):
        self._L.orc_step(self._s, n)

    def set_state(self, h, um, vn):
        """Replace (h, um, vn) and recompute the derived fields (test aid)."""
        a = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (h, um, vn)]
        self._L.orc_set_state(self._s, _dp(a[0]), _dp(a[1]), _dp(a[2]))

    def round_state_f32(self):
        self._L.orc_round_state_f32(self._s)

    def run_until(self, tmax, max_iters=1 << 62):
        return self._L.orc_run_until(self._s, tmax, max_iters)

    @property
    def time(self):
        return self._L.orc_time(self._s)

    @property
    def mstep(self):
        return self._L.orc_mstep(self._s)

    @property
    def dt2_next(self):
        return self._L.orc_dt2_next(self._s)

    def field(self, name):
        out = np.zeros(self.R * self.C)
        assert self._L.orc_get_field(self._s, FIELDS[name], _dp(out)) == 0
        return out.reshape(self.R, self.C)

    def ledger(self):
        out = np.zeros(8)
        self._L.orc_ledger(self._s, _dp(out))
        keys = ("volume", "v0", "v_src", "v_clip", "v_out", "q_out", "q_peak", "t_peak")
        return dict(zip(keys, out))

    def hydrograph(self, cap=1 << 20):
        out = np.zeros(2 * cap)
        n = self._L.orc_hydrograph(self._s, _dp(out), cap)
        n = min(n, cap)
        return out[: 2 * n].reshape(n, 2)

    def bc_acc(self):
        out = np.zeros(max(self.nbc, 1))
        self._L.orc_bc_acc(self._s, _dp(out))
        return out[: self.nbc]
