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


try:  # scipy's pocketfft gives numpy's values bit for bit, on several threads
    import scipy.fft as _sfft

    def _irfft2(a, s):
        return _sfft.irfft2(a, s=s, workers=-1)
except ImportError:  # pragma: no cover
    def _irfft2(a, s):
        return np.fft.irfft2(a, s=s)


def fbm(rows, cols, hurst, amplitude, seed):
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(rows)[:, None]
    kx = np.fft.rfftfreq(cols)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    noise = rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)
    surf = _irfft2(noise * amp, s=(rows, cols))
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


def block_rain(intensity_mm_h, duration_min, n_after=3):
    """rain.dat of a block storm: one interval of the storm's length holding its
    depth in mm, then dry intervals (config 0 writes 50 mm/h for 30 min as
    dtrain = 1800 s, values 25, 0, 0, 0)."""
    dtrain = duration_min * 60.0
    return np.array([intensity_mm_h * duration_min / 60.0] + [0.0] * n_after), dtrain


def ensemble_scenario(k, base_seed=BASE_SEED + 3, side=None, t_end=3600.0, field_cache=None):
    """Scenario k of the batched ensemble: square grid, side = round(128 * 8^U)
    (log-uniform 128..1024) unless given, dx = 10 m; DEM as config 0 with slope
    U(0.2 %, 3 %) and fBm amplitude U(0.5, 5) m; n ~ U(0.02, 0.1); block rain
    U(10, 120) mm/h for U(15, 90) min; outlet at the lowest south-edge cell;
    fixed dt = 0.5 s; 1 h.  Seeded by base_seed + k.  field_cache maps a side
    to a list of unit-amplitude fBm fields to draw from instead of synthesising
    one per scenario (the serving generator uses it: thousands of requests)."""
    rng = np.random.default_rng(base_seed + k)
    if side is None:
        side = int(round(128 * 8.0 ** rng.random()))
    dx = 10.0
    slope = rng.uniform(0.002, 0.03)
    amp = rng.uniform(0.5, 5.0)
    n = rng.uniform(0.02, 0.1)
    inten = rng.uniform(10.0, 120.0)
    dur = rng.uniform(15.0, 90.0)
    if field_cache is not None:
        pool = field_cache[side]
        f = pool[int(rng.integers(len(pool)))] * amp
    else:
        f = fbm(side, side, 0.7, amp, base_seed + k)
    r = np.arange(side, dtype=np.float64)[:, None]
    z = slope * dx * (side - 1 - r) + f
    rain, dtrain = block_rain(inten, dur, n_after=max(3, int(np.ceil(t_end / (dur * 60.0))) + 1))
    spec = CaseSpec(f"ens{k}", z, np.zeros((side, side)), dx, dt=0.5, manning=n, t_end=t_end,
                    rain=rain, dtrain=dtrain)
    spec.outlet = (side - 1, int(np.argmin(z[side - 1])))
    spec.extra = dict(slope=slope, amp=amp, intensity=inten, duration=dur)
    return spec


SERVE_SIDES = (256, 512, 1024, 2048, 4096)
SERVE_PROBS = (0.4, 0.3, 0.15, 0.1, 0.05)


def serve_requests(n, rate=20.0, seed=BASE_SEED + 4, t_end=7200.0, fields_per_side=2,
                   sides=SERVE_SIDES, probs=SERVE_PROBS):
    """The serving stream: Poisson arrivals at `rate` per second, grid side drawn
    from `sides` with `probs`, scenarios as ensemble_scenario (2 h storms).
    Returns (arrival offsets in s, specs).  DEM relief comes from a cache of
    `fields_per_side` seeded unit fBm fields per side, scaled per request."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    chosen = rng.choice(len(sides), size=n, p=np.asarray(probs) / np.sum(probs))
    cache = {}
    for i in sorted(set(chosen.tolist())):
        s = sides[i]
        cache[s] = [fbm(s, s, 0.7, 1.0, seed + 1000 * s + j) for j in range(fields_per_side)]
    specs = [ensemble_scenario(k, base_seed=seed + 17, side=sides[chosen[k]], t_end=t_end,
                               field_cache=cache) for k in range(n)]
    return arrivals, specs


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


# ---------------------------------------------------------------------------
# db/<case>/ tree in the reference's file layout (read by sw2d-efdc)
# ---------------------------------------------------------------------------
_SETUP1 = """#*******************************************************#
# 1-SETUP: SW2D-GPU                                     #
#*******************************************************#
# START TIME: DAY, HOUR, MINUTES, SECONDS #

{t0}

# END TIME: DAY, HOUR, MINUTES, SECONDS #

{t1}

# TIME STEP #

{dt}

# TIME INTERVAL TO SHOW TIME ON SCREEM (SECONDS) #

{dpout}

# TIME INTERVAL TO WRITE THE RESULTS (SECONDS) #

{dkout}

# TEMPORAL RESOLUTION OF FLOW INPUT DATA (SECONDS) #

{dtoq}

# EVAPORATION IN THE LAKE: 1 ACTIVE, 0 INACTIVE #

{evap}

# MINIMUM WATER DEPTH FOR EVAPORATION #

{ev_wl_min}

# PERCENTAGE OF WATER LOST BY INTERCEPTION #

{INT}

# PERCENTAGE OF WATER LOST BY INFILTRATION #

{INF}

# PERCENTAGE OF WATER LOSS IN THE LAKE #

{LWL}

# MANNING COEFFICIENT #

{manning}

# OUTPUT X VELOCITY: 1 ACTIVE, 0 INACTIVE #

{ovx}

# OUTPUT Y VELOCITY: 1 ACTIVE, 0 INACTIVE #

{ovy}

# OUTPUT ELEVATIONS: 1 ACTIVE, 0 INACTIVE #

{oel}

# OUTPUT WATER DEPTH: 1 ACTIVE, 0 INACTIVE #

{odp}

# OUTPUT WATER DEPTH TIME SERIES: 1 ACTIVE, 0 INACTIVE #

{oot}

#*******************************************************#
# 2-SETUP: EFDC-MPI                                     #
#*******************************************************#
# 3D HYDRODYNAMIC MODEL: 1 ACTIVE, 0 INACTIVE #

{efdc}
"""

# the EFDC block: (comment lines, value) in file order
_EFDC = [
    (["#-------------- TEMPORAL CONTROLLERS -------------------#", "# EFDC BASE TIME #"], "basetime"),
    (["# TIME ORIGIN OF RUN #"], 0.0),
    (["# NUMBER OF TIME STEPS PER REFERENCE TIME PERIOD #"], 8640),
    (["#------- ROUGHNESS AND DEPTH PARAMETERS ----------#", "# NUMBER OF VERTICAL LAYERS #"], "nlayers"),
    (["# LOG BDRY LAYER CONST OR VARIABLE ROUGH HEIGHT ADJ IN METERS #"], 0.02),
    (["# LOG BDRY LAYER VARIABLE ROUGHNESS HEIGHT CONVERT TO METERS #"], 1.0),
    (["# MINIMUM DEPTH OF INPUTS DEPTHS IN METERS #"], 0.1),
    (["# ADJUCTMENT TO DEPTH FIELD IN METERS #"], 0.0),
    (["# CONVERTS INPUT DEPTH FIELD TO METERS #"], 1.0),
    (["# DEPTH AT WHICH CELL OR FLOW FACE BECOMES DRY #"], 0.02),
    (["# DEPTH AT WHICH CELL OR FLOW FACE BECOMES WET #"], 0.05),
    (["# ADJUCTMENT TO BOTTOM BED ELEVATION FIELD IN METERS #"], 0.0),
    (["# CONVERTS INPUT BOTTOM BED ELEVATION FIELD TO METERS #"], 1.0),
    (["#----------------- ATMOSPHERIC FORCING -----------------#",
      "# NUMBER OF WIND TIME SERIES (0 SETS WIND TO ZERO) #"], 0),
    (["# NUMBER OF ATMOSPHERIC CONDITION TIME SERIES (0 SETS ALL  ZERO) #"], 0),
    (["#---------- ATMOSPHERIC AND THERMAL PROPERTIES ---------#",
      "# 0 DISTRIBUTE SW SOL RAD OVER WATER COL AND INTO BED, = 1 ALL TO SURF LAYER #"], 1),
    (["# 1000 * EVAPORATIVE TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF #"], 1.5),
    (["# 1000 * CONVECTIVE HEAT TRANSFER COEF, REVC<0 USE WIND SPD DEPD DRAG COEF #"], 1.5),
    (["# FAST SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1./METERS #"], 2.0),
    (["# SLOW SCALE SOLAR SW RADIATION ATTENUATION COEFFCIENT 1./METERS #"], 0.5),
    (["# FRACTION OF SOLSR SW RADIATION ATTENUATED FAST  0<FSWRATF<1 #"], 0.5),
    (["# DEPTH OR THICKNESS OF ACTIVE BED TEMPERATURE LAYER, METERS #"], 1.0),
    (["# INITIAL BED TEMPERATURE #"], 20.0),
    (["# CONVECTIVE HT COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  NO DIM #"], 0.0),
    (["# HEAT TRANS COEFFCIENT BETWEEN BED AND BOTTOM WATER LAYER  M/SEC #"], 0.0),
    (["#----- CONTROLS FOR HORIZONTAL SURFACE ELEVATION OR PRESSURE CONTOURING -----#",
      "# SURFACE ELEVATION OUTPUT #",
      "# 1 TO WRITE FILE FOR SURFACE ELEVATION OR PRESSURE CONTOURING #",
      "# 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD #"], 0),
    (["# NUMBER OF WRITES PER REFERENCE TIME PERIOD #"], 1),
    (["# 1 TO WRITE FILE FOR RESIDUAL SURFACE ELEVATION  CONTOURNG IN HORIZONTAL PLANE #"], 0),
    (["# SURFACE ELEVATION WRITING FORMAT #",
      "# 0 DOES NOT WRITE I,J,X,Y IN surfplt.out and rsurfplt.out FILES #",
      "# 1 WRITES I,J ONLY IN surfplt.out and rsurfplt.out FILES #",
      "# 2 WRITES I,J,X,Y  IN surfplt.out and rsurfplt.out FILES #",
      "# 3 WRITES EFDC EXPLORER BINARY FORMAT FILES #"], 0),
    (["#------ CONTROLS FOR HORIZONTAL PLANE VELOCITY VECTOR PLOTTING ----------#",
      "# HORIZONTAL PLANE VELOCITY OUTPUT #",
      "# 1 TO WRITE FILE FOR VELOCITY PLOTTING IN HORIZONTAL PLANE #",
      "# 2 WRITE ONLY DURING LAST REFERENCE TIME PERIOD #"], 0),
    (["# NUMBER OF WRITES PER REFERENCE TIME PERIOD #"], 1),
    (["# 1 TO WRITE FILE FOR RESIDUAL VELOCITY PLOTTIN IN HORIZONTAL PLANE #"], 0),
    (["# HORIZONTAL VELOCITY WRITING FORMAT #",
      "# 0 DOES NOT WRITE I,J,X,Y IN velplth.out and rvelplth.out FILES #",
      "# 1 WRITES I,J ONLY IN velplth.out and rvelplth.out FILES #",
      "# 2 WRITES I,J,X,Y  IN velplth.out and rvelplth.out FILES #",
      "# 3 WRITES EFDC EXPLORER BINARY FORMAT FILES #"], 0),
    (["#--------------- CONTROLS FOR WRITING TO TIME SERIES FILES --------------#",
      "# WRITE SCENARIO FOR THIS LOCATION #"], 1),
    (["# 1 FOR TIME SERIES OF SURFACE ELEVATION #"], 1),
    (["# 1 FOR TIME SERIES OF TRANSPORTED CONCENTRATION VARIABLES #"], 0),
    (["# 1 FOR TIME SERIES OF EDDY VISCOSITY AND DIFFUSIVITY #"], 0),
    (["# 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL VELOCITY #"], 0),
    (["# 1 FOR TIME SERIES OF EXTERNAL MODE HORIZONTAL TRANSPORT #"], 0),
    (["# 1 FOR TIME SERIES OF HORIZONTAL VELOCITY IN EVERY LAYER #"], 0),
    (["#---------------------- SHOW.INP PARAMETERS ---------------------------#", "# NSTYPE #"], 1),
    (["# NSHOWR #"], 1), (["# ISHOWC #"], 1), (["# JSHOWC #"], 1), (["# NSHFREQ #"], 1),
    (["# ZSSMIN #"], 0), (["# ZSSMAX #"], 1), (["# SSALMAX #"], 0),
    (["#----------------- DOMAIN DECOMPOSITION PARAMETERS --------------------#",
      "# NUMBER OF PROCESSORS #"], 1),
    (["# LOCALLY OPTIMAL RECTILINEAR PARTITION (LORP) TYPE #", "# 1 TYPE h #", "# 2 TYPE v #",
      "# 3 TYPE r #"], 1),
    (["#---------------- CHOOSE TO INCLUDE IN THE SIMULATION -----------------#",
      "# DYE SIMULATION: 1 ACTIVE, 0 INACTIVE #"], 0),
    (["# TEMPERATURE SIMULATION: 1 ACTIVE, 0 INACTIVE #"], 0),
    (["# SALINITY SIMULATION: 1 ACTIVE, 0 INACTIVE #"], 0),
    (["#----------------------- INITIAL CONDITIONS ---------------------------#",
      "# INITIAL DYE CONCENTRATION IN THE LAKE #"], 0.0),
    (["# INITIAL SALINITY IN THE LAKE #"], 0.0),
    (["# WATER TEMPERATURE AT THE BOTTOM OF THE LAKE #"], 20.0),
    (["# WATER TEMPERATURE ON THE LAKE SURFACE #"], 22.0),
    (["#---- CATCHMENT-LAKE BOUNDARIES (WATER INPUTS FROM THE CATCHMENT) -----#",
      "# DYE CONCENTRATION #"], 0.0),
    (["# SALINITY #"], 0.0),
]

# names of the EFDC block's values, in _EFDC order (the keys of write_db's
# efdc_values)
EFDC_NAMES = ["basetime", "TBEGIN", "NTSPTC", "nlayers", "ZBRADJ", "ZBRCVRT", "HMIN", "HADJ",
              "HCVRT", "HDRY", "HWET", "BELADJ", "BELCVRT", "NWSER", "NASER", "IASWRAD", "REVC",
              "RCHC", "SWRATNF", "SWRATNS", "FSWRATF", "DABEDT", "TBEDIT", "HTBED1", "HTBED2",
              "ISPPH", "NPPPH", "ISRPPH", "IPPHXY", "ISVPH", "NPVPH", "ISRVPH", "IVPHXY", "NTSSSS",
              "MTSP", "MTSC", "MTSA", "MTSUE", "MTSUT", "MTSU", "NSTYPE", "NSHOWR", "ISHOWC",
              "JSHOWC", "NSHFREQ", "ZSSMIN", "ZSSMAX", "SSALMAX", "nprocessors", "lorp", "dye_on",
              "temp_on", "salt_on", "DYE_INIT", "SALT_INIT", "TEMP_BOT", "TEMP_SURF", "DYE_IN",
              "SALT_IN"]
assert len(EFDC_NAMES) == len(_EFDC)


def efdc_value(name, efdc_values=None, basetime=86400.0, nlayers=1):
    """The value write_db puts in the EFDC block for `name`."""
    i = EFDC_NAMES.index(name)
    if efdc_values and name in efdc_values:
        return efdc_values[name]
    v = _EFDC[i][1]
    return {"basetime": basetime, "nlayers": nlayers}.get(v, v) if isinstance(v, str) else v


def _hms(t):
    d, r = divmod(int(round(t)), 86400)
    h, r = divmod(r, 3600)
    m, s = divmod(r, 60)
    return f"{d} {h} {m} {s}"


def _esri(path, a, cellsize, nodata, fmt, xll=0.0, yll=0.0):
    R, C = a.shape
    with open(path, "w") as f:
        f.write(f"ncols         {C}\nnrows         {R}\nxllcorner     {float(xll)!r}\n"
                f"yllcorner     {float(yll)!r}\ncellsize      {float(cellsize)!r}\n"
                f"NODATA_value  {float(nodata)!r}\n")
        for row in a:
            f.write(" ".join(fmt(v) for v in row) + "\n")


# grid_make_input.dat: the grid generator's controls, in the file's schema
GRID_MAKE_INPUT = """SOLUTION ITERATIONS AND SMOOTHING INFORMATION
ITRXM: 150
ITRHM: 160
ITRKM: 170
ITRGM: 180
NDEPSM: 2000
NDEPSMF: 1
DEPMIN: 0.5
DDATADJ: 0
PARAMETERS AND CONVERGENCE CRITERIA
RPX: 1.7
RPK: 1.6
RPH: 1.5
RSQXM: 1.E-10
RSQKM: 2.E-11
RSQKIM: 3.E-12
RSQHM: 4.E-13
RSQHIM: 5.E-14
RSQHJM: 6.E-15
COORDINATE SHIFT PARAMETERS
XSHIFT: 10.
YSHIFT: 20.
HSCALE: 1.
RKJDKI: 1.
ANGORO: 4.0
INTERPOLATION SWITCHES
ISIRKI: 1
JSIRKI: 0
ISIHIHJ: 1
JSIHIHJ: 0
DEPTH INTERPOLATION SWITCHES
ISIDEP: 1
CDEP: 2.
RADM: 3.
ISIDPTYP: 2
SURFELV: 9.5
ISVEG: 0
NVEGDAT: 0
NVEGTYP: 0
ZROUGH: 0.0005
VERBOSE: 1"""


def efdc_card_template():
    """A small EFDC.INP card template (CRLF lines, as EFDC's own): cards C1 to
    C91 with C11A and C71A, each a title line, a comment, a header line, one
    data line and a dashed line."""
    names = []
    for k in range(1, 92):
        names.append(f"C{k}")
        if k in (11, 71):
            names.append(f"C{k}A")
    out = ["C ** EFDC.INP card template  ", "C **  "]
    for i, nm in enumerate(names):
        out += [f"{nm} TITLE OF CARD {nm}  ", "*  ", f"*  {nm}X: A VALUE OF CARD {nm}  ",
                f"{nm}  {nm}X  {nm}Y  ", f"    {i}    {i + 1}  ", "-" * 79 + "  "]
    return ("\r\n".join(out) + "\r\n").encode()


def write_db(spec, root, name, *, dkout=600.0, dpout=600.0, probes=(), sources=None,
             efdc=False, basetime=86400.0, nlayers=1, evaporation=None, INT=0.0, INF=0.0,
             LWL=0.0, EV_WL_min=0.0, outputs=(1, 1, 1, 1, 1), efdc_values=None, met=None,
             efdc_points=((15.0, 25.0),), tctmsr=1.0, efdc_template=None, grid_make=None):
    """Write db/dir.txt and db/<name>/ for a fixed-step spec in the reference's
    layout.  probes: (x, y) metres for coord_out.dat; sources: (xs, ys, Q, dtoq);
    evaporation: (albedo, T[], Rg[]) with one value per rain interval.
    efdc_values: EFDC block values by EFDC_NAMES name; met: (patm[], relh[],
    wind[(speed, direction)]) for atm_press.dat, rel_humidity.dat and
    wind_velocities.dat (constant series when None); efdc_points, tctmsr:
    efdc_coord_out.dat; efdc_template: bytes of EFDC/input/template/EFDC.INP
    (efdc_card_template() when None); grid_make: text of grid_make_input.dat
    (GRID_MAKE_INPUT when None)."""
    import os
    base = os.path.join(root, "db", name)
    for d in ("SW2D/input", "SW2D/output", "EFDC/input", "EFDC/output"):
        os.makedirs(os.path.join(base, d), exist_ok=True)
    with open(os.path.join(root, "db", "dir.txt"), "w") as f:
        f.write(f"dir_number: 1\n{name}\n")
    dtoq = sources[3] if sources else 3600.0
    ovx, ovy, oel, odp, oot = outputs
    txt = _SETUP1.format(t0="0 0 0 0", t1=_hms(spec.t_end), dt=repr(spec.dt), dpout=repr(dpout),
                         dkout=repr(dkout), dtoq=repr(dtoq),
                         evap=1 if evaporation else 0, ev_wl_min=repr(EV_WL_min), INT=repr(INT),
                         INF=repr(INF), LWL=repr(LWL), manning=repr(spec.manning), ovx=ovx,
                         ovy=ovy, oel=oel, odp=odp, oot=oot if probes else 0,
                         efdc=1 if efdc else 0)
    if efdc:
        lines = []
        for (comments, v), nm in zip(_EFDC, EFDC_NAMES):
            lines += comments
            val = efdc_value(nm, efdc_values, basetime, nlayers)
            lines += ["", repr(float(val)) if isinstance(val, float) else str(val), ""]
        txt += "\n" + "\n".join(lines) + "\n"
    with open(os.path.join(base, "setup.dat"), "w") as f:
        f.write(txt)
    inp = os.path.join(base, "SW2D", "input")
    _esri(os.path.join(inp, "dem.dat"), spec.z, spec.cellsize, spec.nodata,
          lambda v: repr(float(v)))
    with open(os.path.join(inp, "initial_condition.dat"), "w") as f:
        for row in spec.h0:
            f.write(" ".join(repr(float(v)) for v in row) + "\n")
    xs, ys, Q = (sources[0], sources[1], np.asarray(sources[2]).reshape(-1, len(sources[0]))) \
        if sources else ([], [], np.zeros((1, 0)))
    with open(os.path.join(inp, "coord_source_sink.dat"), "w") as f:
        f.write(f"nst {len(xs)}\n" + "".join(f"{x!r} {y!r}\n" for x, y in zip(xs, ys)))
    with open(os.path.join(inp, "Q_source_sink.dat"), "w") as f:
        f.write("flow (m3/s)\n" + "\n".join(" ".join(repr(float(v)) for v in row) for row in Q))
    rain = spec.rain if spec.rain is not None else np.zeros(int(spec.t_end // spec.dtrain) + 2)
    with open(os.path.join(inp, "rain.dat"), "w") as f:
        # no newline after the last row: the reader takes (newlines - 1) values
        f.write(f"{spec.dtrain!r}\nTime (hour) Rainfall (mm)\n" +
                "\n".join(f"{i} {float(v)!r}" for i, v in enumerate(rain)))
    if probes:
        with open(os.path.join(inp, "coord_out.dat"), "w") as f:
            f.write(f"n_out {len(probes)}\n" + "".join(f"{x!r} {y!r}\n" for x, y in probes))
    if evaporation:
        albedo, T, Rg = evaporation
        with open(os.path.join(inp, "temperatures.dat"), "w") as f:
            f.write("Temperature\n" + "\n".join(repr(float(v)) for v in T) + "\n")
        with open(os.path.join(inp, "solar_radiation.dat"), "w") as f:
            f.write(f"albedo: {albedo!r}\nRadiation\n" + "\n".join(repr(float(v)) for v in Rg) + "\n")
    if efdc:
        mask = spec.lake_mask if spec.lake_mask is not None else np.zeros(spec.z.shape, np.int32)
        _esri(os.path.join(inp, "mask.dat"), mask.astype(np.int64), spec.cellsize, -9999,
              lambda v: str(int(v)))
        n = len(rain) + 2
        patm, relh, wind = met if met else ([1000.0] * n, [0.8] * n, [(0.0, 0.0)] * n)
        with open(os.path.join(inp, "atm_press.dat"), "w") as f:
            f.write("atm_pressure(millibar)\n" + "".join(f"{float(v)!r}\n" for v in patm))
        with open(os.path.join(inp, "rel_humidity.dat"), "w") as f:
            f.write("rel_humidity(percent)\n" + "".join(f"{float(v)!r}\n" for v in relh))
        with open(os.path.join(inp, "wind_velocities.dat"), "w") as f:
            f.write("".join(f"{float(a)!r}\t{float(b)!r}\n" for a, b in wind))
        with open(os.path.join(inp, "efdc_coord_out.dat"), "w") as f:
            f.write(f"time_conversion: {float(tctmsr)!r}\nnumber_t_s_out: {len(efdc_points)}\n" +
                    "".join(f"{float(x)!r} {float(y)!r}\n" for x, y in efdc_points))
        tdir = os.path.join(base, "EFDC", "input", "template")
        os.makedirs(tdir, exist_ok=True)
        with open(os.path.join(tdir, "EFDC.INP"), "wb") as f:
            f.write(efdc_template if efdc_template is not None else efdc_card_template())
        with open(os.path.join(base, "grid_make_input.dat"), "w") as f:
            f.write(grid_make if grid_make is not None else GRID_MAKE_INPUT)
    return base
