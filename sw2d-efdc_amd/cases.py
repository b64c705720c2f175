"""Synthetic SW2D cases (DEMs, depths, forcing) for the benchmark configurations.

All inputs are generated, seeded and documented here; nothing is downloaded.
Conventions follow the reference grid: row 0 is the north row, "south" is the
last row, cellsize dx = dy, Manning n per cell or scalar.

fBm: spectral synthesis (Voss/Saupe): complex white noise in the Fourier
domain scaled by |k|^-(H+1) (power spectrum |k|^-(2H+2)), inverse FFT, real
part, normalised so that max - min equals the stated amplitude.
Random streams: numpy PCG64 seeded with base_seed + scenario index.
"""
from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 20261202


def fbm(rows, cols, hurst, amplitude, seed):
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(rows)[:, None]
    kx = np.fft.rfftfreq(cols)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    noise = rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)
    surf = np.fft.irfft2(noise * amp, s=(rows, cols))
    span = surf.max() - surf.min()
    if span > 0:
        surf = (surf - surf.min()) / span * amplitude
    return surf


@dataclass
class CaseSpec:
    """Everything sw2d_create and the set_* calls need for one case."""
    name: str
    z: np.ndarray
    h0: np.ndarray
    cellsize: float
    dt: float                       # half step (setup.dat TIME STEP); dt2 max when adaptive
    manning: float
    t_end: float
    adaptive: bool = False
    courant: float = 0.5
    rain: np.ndarray = None         # mm per dtrain interval
    dtrain: float = 3600.0
    nclass: np.ndarray = None       # uint8 class per cell
    ntable: np.ndarray = None
    storm: dict = None
    outlet: tuple = None
    lake_mask: np.ndarray = None
    nodata: float = -9999.0
    extra: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.z.shape[0] * self.z.shape[1]


def config0(rows=256, cols=256, seed=BASE_SEED, outlet=False):
    """Plain FP64 reference case: 256x256, dx = 10 m, 1 % plane sloping south +
    fBm (H = 0.7, 2 m), n = 0.03, dry start, rain 50 mm/h for 30 min then 30 min
    recession (rain.dat: dtrain 1800 s, values 25, 0, 0, 0 mm), dt = 0.5 s,
    1 h = 3600 iterations.  outlet=True adds the free-outflow cell at the
    lowest cell of the south edge."""
    dx = 10.0
    r = np.arange(rows, dtype=np.float64)[:, None]
    z = 0.01 * dx * (rows - 1 - r) + fbm(rows, cols, 0.7, 2.0, seed)
    z = np.broadcast_to(z, (rows, cols)).copy()
    spec = CaseSpec("config0", z, np.zeros((rows, cols)), dx, dt=0.5, manning=0.03, t_end=3600.0,
                    rain=np.array([25.0, 0.0, 0.0, 0.0]), dtrain=1800.0)
    if outlet:
        spec.outlet = (rows - 1, int(np.argmin(z[rows - 1])))
    return spec


def config1(rows=2048, cols=2048, seed=BASE_SEED + 1, t_end=7200.0):
    """Plain FP64 drop-in at real size: 2048x2048, dx = 5 m, 0.5 % regional slope
    to the south + 3 meandering Gaussian-valley channels (FWHM 20 m, depth 2 m)
    + fBm 3 m + a Gaussian lake pit (depth 8 m, sigma = radius/2 = 200 m)
    prefilled 5 m above its lowest point; lake cells (mask 1) are the prefilled
    ones and their catchment rim gives the EFDC inflow records.  n = 0.02 in
    channel cells (within 10 m of a channel axis), 0.08 elsewhere.  Storm:
    Gaussian footprint sigma 1.5 km, peak 80 mm/h, centre starting at the west
    edge mid-height and moving east at 5 m/s for 1 h.  CFL-adaptive step
    (Courant 0.5, half step at most 0.5 s); 2 h simulated."""
    dx = 5.0
    rng = np.random.default_rng(seed)
    y = (np.arange(rows, dtype=np.float64)[:, None] + 0.5) * dx
    x = (np.arange(cols, dtype=np.float64)[None, :] + 0.5) * dx
    Ly, Lx = rows * dx, cols * dx
    z = 0.005 * (Ly - y) + fbm(rows, cols, 0.7, 3.0, seed)
    sig_c = 20.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    chan = np.zeros((rows, cols), dtype=bool)
    for i, xf in enumerate((0.25, 0.5, 0.75)):
        amp = 0.02 * Lx * (1 + 0.5 * rng.random())
        lam = 0.3 * Ly * (1 + rng.random())
        xc = xf * Lx + amp * np.sin(2 * np.pi * y / lam + rng.random() * 2 * np.pi)
        d = x - xc
        z = z - 2.0 * np.exp(-(d * d) / (2 * sig_c * sig_c))
        chan |= np.abs(d) <= 10.0
    # lake pit in the southern half, between two channels
    px, py, sig_p = 0.62 * Lx, 0.7 * Ly, 200.0
    d2 = (x - px) ** 2 + (y - py) ** 2
    z = z - 8.0 * np.exp(-d2 / (2 * sig_p * sig_p))
    near = d2 <= (3 * sig_p) ** 2
    level = z[near].min() + 5.0
    h0 = np.where(near & (z < level), level - z, 0.0)
    lake = (h0 > 0).astype(np.int32)
    ncls = np.where(chan, 0, 1).astype(np.uint8)
    spec = CaseSpec("config1", z, h0, dx, dt=0.5, manning=0.08, t_end=t_end, adaptive=True,
                    courant=0.5, nclass=ncls, ntable=np.array([0.02, 0.08]),
                    storm=dict(peak=80.0 / 3.6e6, x0=0.0, y0=0.5 * Ly, vx=5.0, vy=0.0,
                               sigma=1500.0, t0=0.0, t1=3600.0),
                    lake_mask=lake)
    return spec


def config2(rows=4096, cols=4096, seed=BASE_SEED + 2, t_end=1800.0):
    """Precision-form pond break: 4096x4096, dx = 2 m, flat floodplain + fBm
    0.5 m; a 400 m x 400 m (200 x 200 cells) reservoir in the north-west
    quarter held 4 m deep (free surface 4 m above the reservoir's mean bed) and
    released at t = 0; dry downstream bed, n = 0.035, no rain, adaptive step,
    30 min."""
    dx = 2.0
    z = fbm(rows, cols, 0.7, 0.5, seed)
    r0, c0 = rows // 8, cols // 8
    res = np.zeros((rows, cols), dtype=bool)
    res[r0:r0 + 200, c0:c0 + 200] = True
    level = z[res].mean() + 4.0
    h0 = np.where(res, np.maximum(level - z, 0.0), 0.0)
    return CaseSpec("config2", z, h0, dx, dt=0.5, manning=0.035, t_end=t_end, adaptive=True,
                    courant=0.5)


def apply(case_obj, spec):
    """Configure a sw2d Case (or the CPU oracle) with the forcing of a spec."""
    if spec.rain is not None:
        case_obj.set_rain(spec.rain, spec.dtrain)
    if spec.nclass is not None:
        if hasattr(case_obj, "set_manning_map") and case_obj.__class__.__name__ == "Oracle":
            case_obj.set_manning_map(spec.ntable[spec.nclass])
        else:
            case_obj.set_manning_map(spec.nclass, spec.ntable)
    if spec.storm is not None:
        s = spec.storm
        case_obj.set_storm(s["peak"], s["x0"], s["y0"], s["vx"], s["vy"], s["sigma"], s["t0"],
                           s["t1"])
    if spec.outlet is not None:
        case_obj.set_outlet(*spec.outlet)
    if spec.adaptive:
        if case_obj.__class__.__name__ == "Oracle":
            case_obj.set_adaptive(spec.courant, 2.0 * spec.dt)
    return case_obj
