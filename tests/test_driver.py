"""The sw2d-efdc drop-in driver: the reference's db/ file contract.

CPU tests parse generated cases (--parse-only) and check the EFDC inputs
known before a run (--efdc-inputs).  The GPU test runs a whole coupled case
and compares every output file with the same files built from the CPU FP64
oracle, formatted with the reference's printf formats (SW2D-EFDC.cu:4958-5072
for VTK and WaterLevel.txt, 2499-2610 for QSER.INP, 2026-2498 for the
constituent series).
"""
import json
import os
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
# SW2D_DRIVER selects another build of the driver (the sanitizer build)
DRIVER = pathlib.Path(os.environ.get("SW2D_DRIVER") or ROOT / "sw2d-efdc_amd" / "sw2d-efdc")


def small_case(cases, R=24, C=37, t_end=480.0):
    spec = cases.config0(rows=R, cols=C, seed=11)
    spec.t_end = t_end
    spec.rain = np.array([60.0, 30.0, 30.0, 0.0, 0.0])
    spec.dtrain = 120.0
    mask = np.zeros((R, C), np.int32)
    mask[R // 2:R // 2 + 4, C // 3:C // 3 + 6] = 1
    spec.lake_mask = mask
    return spec


CASE_KW = dict(dkout=120.0, dpout=60.0, probes=[(55.0, 105.0), (125.0, 35.0), (305.0, 205.0)],
               sources=([105.0], [155.0], [[0.5], [1.0], [0.2], [0.0]], 150.0), efdc=True,
               nlayers=2, basetime=3600.0,
               evaporation=(0.2, [20.0, 21.0, 22.0, 23.0, 24.0], [300.0, 400, 500, 600, 700]),
               INT=0.1, INF=0.05, LWL=0.2, EV_WL_min=0.01,
               efdc_values=dict(dye_on=1, temp_on=1, salt_on=1, DYE_IN=10.0, SALT_IN=15.0,
                                DYE_INIT=0.5, SALT_INIT=25.0, TEMP_BOT=5.0, TEMP_SURF=23.0,
                                NWSER=1, NSTYPE=3, NSHOWR=7200, ISHOWC=4, JSHOWC=5, NSHFREQ=720,
                                ZSSMIN=-1, ZSSMAX=1300, SSALMAX=1000))


def run_driver(root, *args, check=True):
    r = subprocess.run([str(DRIVER), "--root", str(root), *args], capture_output=True, text=True,
                       timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"sw2d-efdc failed ({r.returncode}): {r.stdout[-2000:]}{r.stderr[-2000:]}")
    return r


def test_parse_only_roundtrip(cases, tmp_path):
    spec = small_case(cases)
    cases.write_db(spec, tmp_path, "C1", **CASE_KW)
    r = run_driver(tmp_path, "--parse-only")
    j = json.loads(r.stdout.strip().splitlines()[-1])
    assert (j["rows"], j["cols"], j["cellsize"], j["nodata"]) == (24, 37, 10.0, -9999.0)
    assert (j["dt"], j["dkout"], j["dpout"], j["dtoq"]) == (0.5, 120.0, 60.0, 150.0)
    assert (j["manning"], j["INT"], j["INF"], j["LWL"], j["EV_WL_min"]) == (0.03, 0.1, 0.05, 0.2, 0.01)
    assert (j["efdc"], j["efdc_basetime"], j["nlayers"], j["evaporation"]) == (1, 3600.0, 2, 1)
    assert (j["nst"], j["nq"], j["nrain"], j["dtrain"], j["n_out"]) == (1, 4, 5, 120.0, 3)
    assert j["rain_sum"] == 120.0 and j["timmax"] == 480.0 and j["time0"] == 0.0
    assert j["z_sum"] == pytest.approx(spec.z.sum(), rel=1e-14)
    assert j["mask_lake"] == int(spec.lake_mask.sum())


def test_setup_crlf_and_whitespace(cases, tmp_path):
    """setup.dat is matched like fscanf: any whitespace run (CRLF included)
    between the comment lines and values is accepted."""
    spec = small_case(cases)
    base = cases.write_db(spec, tmp_path, "C2", **CASE_KW)
    p = os.path.join(base, "setup.dat")
    txt = open(p).read().replace("\n", "   \r\n")
    open(p, "w", newline="").write(txt)
    j = json.loads(run_driver(tmp_path, "--parse-only").stdout.strip().splitlines()[-1])
    assert (j["dt"], j["manning"], j["efdc_basetime"], j["nlayers"]) == (0.5, 0.03, 3600.0, 2)


# ---- EFDC series and initial-condition inputs (formats of 2026-2498, 2612-2765) ----

def c_round_half_away(v):
    return float(np.sign(v) * np.floor(abs(v) + 0.5))


def initial_field_text(mask, nlayers, value):
    """DYE/SALT/TEMP.INP: lake cells south row first, numbered in that order;
    the number's width is round(1 + ln(n)/ln(10)) - 1 for n lake cells."""
    R, C = mask.shape
    n = int((mask == 1).sum())
    width = int(c_round_half_away(1 + np.log(n) / np.log(10)) - 1)
    out = ["C", "C", "C", "C", "1"]
    cell = 0
    for jr in range(R):
        r = R - 1 - jr
        for i in range(C):
            if mask[r, i] != 1:
                continue
            cell += 1
            vals = [value(r, i, l) for l in range(nlayers)]
            out.append("%*d     %d     %d     " % (width, cell, i + 1, jr + 1)
                       + "    ".join("%f" % v for v in vals))
    return "\n".join(out) + "\n"


def series_text(kind, times, nrec, value, nlayers, basetime):
    k = kind
    hdr = [f"C **, {k}SER.INP Time Series FILE", "C **", "C **",
           f"C **  InType M{k}SER(NS) TC{k}SER(NS) TA{k}SER(NS) RMULADJ(NS) ADDADJ(NS)", "C **",
           f"C **  IF InType.EQ.1 THEN READ DEPTH WEIGHTS AND SINGLE VALUE OF {k}SER",
           f"C **                      ELSE READ A VALUE OF {k}SER FOR EACH LAYER", "C **",
           "C **  InType = 1 Structure", "C **  WKQ(K), K = 1, KC",
           f"C **  T{k}SER(M, NS)    {k}SER(M, 1, NS) !(M{k}SER(NS) PAIRS FOR NS = 1, N{k}SER SERIES)",
           "C **", "C **  InType = 0 Structure",
           f"C **  T{k}SER(M, NS)    ({k}SER(M, K, NS), K = 1, KC) !(M{k}SER(NS) PAIRS)", "C **"]
    lines = list(hdr)
    edg = 0
    ntq = len(times)
    for label, n in zip(("North", "South", "East", "West"), nrec):
        for _ in range(n):
            edg += 1
            lines.append("      0 %d %f 0 1 0  ' *** %s_edge_%d" % (ntq, basetime, label, edg))
            for t in range(ntq):
                lines.append("      %f " % times[t] + " ".join(["%f" % value(t)] * nlayers))
    return "\n".join(lines) + "\n"


def test_efdc_inputs_written_before_the_run(cases, po, tmp_path):
    """--efdc-inputs writes WSER, ASER, SHOW and the DYE/SALT/TEMP initial
    fields from the case files alone (no GPU): values and layout against a
    restatement of the reference's writers (wind pairs and meteorological
    series per data interval, efdc_ntimes = run length / dt / lkout rows)."""
    spec = small_case(cases)
    rng = np.random.default_rng(5)
    n = len(spec.rain) + 2
    met = (1000.0 + rng.uniform(-20, 20, n), rng.uniform(0.3, 0.9, n),
           [(a, b) for a, b in zip(rng.uniform(0, 8, n), rng.uniform(0, 360, n))])
    kw = dict(CASE_KW, nlayers=3)
    base = pathlib.Path(cases.write_db(spec, tmp_path, "C6", met=met, **kw))
    run_driver(tmp_path, "--efdc-inputs")
    inp = base / "EFDC" / "input"
    ev = kw["efdc_values"]
    nl, basetime, dtrain = 3, kw["basetime"], spec.dtrain
    ntime = int((spec.t_end / spec.dt) / round(kw["dkout"] / spec.dt))
    mask = spec.lake_mask
    assert (inp / "DYE.INP").read_text() == initial_field_text(mask, nl, lambda r, i, l: ev["DYE_INIT"])
    assert (inp / "SALT.INP").read_text() == initial_field_text(mask, nl, lambda r, i, l: ev["SALT_INIT"])
    tb, ts = ev["TEMP_BOT"], ev["TEMP_SURF"]
    temp = lambda r, i, l: tb + (ts - tb) / (1.0 + (spec.h0[r, i] / nl) * ((nl - 1) - l))
    assert (inp / "TEMP.INP").read_text() == initial_field_text(mask, nl, temp)
    # WSER: 17 comment lines, the series header, then (time in base units, speed, direction)
    w = (inp / "WSER.INP").read_text().splitlines()
    assert len(w) == 17 + 1 + ntime and all(l.startswith(("C **", "#")) for l in w[:17])
    assert w[17] == "      %d %f %f %f %d" % (ntime, basetime, 0.0, 1.0, 1)
    for t in range(ntime):
        assert w[18 + t] == "      %f %f %f" % (t * dtrain / basetime, *met[2][t])
    # ASER: 36 comment lines, two parameter lines, then 7 columns per interval
    a = (inp / "ASER.INP").read_text().splitlines()
    assert len(a) == 36 + 2 + ntime and all(l.startswith("#") for l in a[:36])
    cvt = 1 / (dtrain * 1000.0)
    assert a[36] == "      %d %f %f %d %E %E %f %f" % (ntime, basetime, 1.0, 1, cvt, cvt, 1.0, 1.0)
    th = [cases.efdc_value(k, ev) for k in ("IASWRAD", "REVC", "RCHC", "SWRATNF", "SWRATNS",
                                            "FSWRATF", "DABEDT", "TBEDIT", "HTBED1", "HTBED2")]
    assert a[37] == "      %d %f %f %f %f %E %f %f %E %E" % (int(th[0]), *th[1:])
    albedo, T, Rg = kw["evaporation"]
    evap = po.evaporation(albedo, T, Rg, dtrain)
    for t in range(ntime):
        row = [met[0][t], T[t], met[1][t], (1 - kw["LWL"]) * spec.rain[t], evap[t], Rg[t], 0.0]
        assert a[38 + t] == "      %f " % (t * dtrain / basetime) + " ".join("%f" % v for v in row), t
    sh = (inp / "SHOW.INP").read_text().splitlines()
    assert sh[6:] == ["    3    7200    4    5    720", "    -1    1300    1000"]
    assert len(sh) == 8 and all(l.startswith("*") for l in sh[:6])


def test_efdc_inputs_short_series_are_reported(cases, tmp_path):
    """A wind file shorter than the run's EFDC record count ends the run with
    a message (the reference reads past the end of its array)."""
    spec = small_case(cases)
    n = len(spec.rain) + 2
    met = ([1000.0] * n, [0.5] * n, [(1.0, 90.0)] * 2)
    cases.write_db(spec, tmp_path, "C7", met=met, **CASE_KW)
    r = run_driver(tmp_path, "--efdc-inputs", check=False)
    assert r.returncode == 1 and "wind_velocities.dat holds 2 values" in r.stdout


def test_missing_input_is_reported(cases, tmp_path):
    spec = small_case(cases)
    base = cases.write_db(spec, tmp_path, "C3", **CASE_KW)
    os.remove(os.path.join(base, "SW2D", "input", "dem.dat"))
    r = run_driver(tmp_path, "--parse-only", check=False)
    assert r.returncode != 0 and "unknown file - dem.dat" in r.stdout


# ---- expected files from the oracle (formats of the reference's printf calls) ----

def vtk_text(R, C, dx, depth, vv, uu, z):
    out = ["# vtk DataFile Version 2.0", "Brazil", "ASCII", "DATASET STRUCTURED_POINTS",
           f"DIMENSIONS {C} {R} 1", "ASPECT_RATIO %f %f %f" % (dx, dx, 1.0),
           "ORIGIN %f %f %f" % (0.0, 0.0, 0.0), f"POINT_DATA {C * R}"]
    for name, a in (("depth", depth), ("x_velocity", vv), ("y_velocity", -uu),
                    ("Elevations", z)):
        out += [f"SCALARS {name} float 1", "LOOKUP_TABLE default"]
        out += ["%f" % v for v in a[::-1].ravel()]
    return "\n".join(out) + "\n"


def c_round(v):
    """C round(): halves away from zero (Python's round() goes to even)."""
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


def water_level_line(R, C, dx, probes, h):
    idx = []
    for x, y in probes:
        outx = c_round(abs(0.0 - x / dx))
        outy = R - c_round(abs(0.0 - y / dx))
        idx.append(outy * C - (C - (outx + 1)))
    vals = [" %f" % h.ravel()[i] for i in sorted(idx)]
    return "".join(vals) + "\n"


def qser_text(hist, nlayers, basetime):
    hdr = ["C **, qser.inp Time Series FILE", "C **",
           "C **  InType MQSER(NS) TCQSER(NS) TAQSER(NS) RMULADJ(NS) ADDADJ(NS) ICHGQS", "C **",
           "C **  IF InType.EQ.1 THEN READ DEPTH WEIGHTS AND SINGLE VALUE OF QSER",
           "C **                      ELSE READ A VALUE OF QSER FOR EACH LAYER", "C **",
           "C **  InType = 1 Structure", "C **  WKQ(K), K = 1, KC",
           "C **  TQSER(M, NS)    QSER(M, 1, NS)          !(MQSER(NS) PAIRS FOR NS = 1, NQSER SERIES)",
           "C **", "C **  InType = 0 Structure",
           "C **  TQSER(M, NS)    (QSER(M, K, NS), K = 1, KC) !(MQSER(NS) PAIRS)", "C **"]
    lines = list(hdr)
    tq = len(hist) - 1
    edg = 0
    for edge, label in (("north", "North"), ("south", "South"), ("east", "East"), ("west", "West")):
        n = len(hist[0][edge]["q"])
        for b in range(n):
            edg += 1
            lines.append("      0 %d %f 0 1 0 0  ' *** %s_edge_%d" % (tq + 1, basetime, label, edg))
            for t in range(tq + 1):
                v = hist[t][edge]["q"][b] / nlayers
                lines.append("      %f " % hist[t]["time"] + " ".join(["%f" % v] * nlayers))
    return "\n".join(lines) + "\n"


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["fused", "plain"])
def test_driver_outputs_match_oracle(cases, po, tmp_path, kernel):
    """Every output file of a coupled run, with the fused kernel and with the
    reference's unfused sequence, equals the files built from the oracle."""
    spec = small_case(cases)
    base = cases.write_db(spec, tmp_path, "C4", **CASE_KW)
    run_driver(tmp_path, "--kernel", kernel)
    R, C, dx = spec.z.shape[0], spec.z.shape[1], spec.cellsize
    z = po.lake_catchment_smooth(spec.lake_mask, spec.z)
    o = po.Oracle(z, spec.h0, dx, dt=spec.dt, manning=spec.manning, INT=0.1, INF=0.05, LWL=0.2,
                  EV_WL_min=0.01)
    o.set_rain(spec.rain, spec.dtrain)
    xs, ys, Q, dtoq = CASE_KW["sources"]
    o.set_sources([po.source_index(R, C, 0.0, 0.0, dx, xs[0], ys[0])], Q, dtoq)
    albedo, T, Rg = CASE_KW["evaporation"]
    o.set_evaporation(po.evaporation(albedo, T, Rg, spec.dtrain))
    dt, lkout = spec.dt, round(CASE_KW["dkout"] / spec.dt)
    time, mstep, out0, tq = 0.0, 0, -1, -1
    uua_l, vva_l = np.zeros((R, C)), np.zeros((R, C))
    hist, wl = [], []
    outdir = pathlib.Path(base) / "SW2D" / "output"
    nfiles = 0
    while time + dt <= spec.t_end:
        out0 += 1
        o.step(1)
        time += dt
        mstep += 1
        time += dt
        mstep += 1
        if not (mstep % lkout == 0 or out0 == 0):
            continue
        tq += 1
        tempo = "0" if out0 == 0 else str(int(round(mstep * dt / CASE_KW["dkout"])))
        h = spec.h0 if tq == 0 else o.field("h")
        rec = po.couple_bc(spec.lake_mask, h, vva_l, uua_l, dx, dx)
        rec["time"] = 0.0 if tq == 0 else mstep * dt / CASE_KW["basetime"]
        hist.append(rec)
        uua_l, vva_l = o.field("uua"), o.field("vva")
        depth = spec.h0 if out0 == 0 else o.field("h")
        wl.append(water_level_line(R, C, dx, CASE_KW["probes"], depth))
        exp = vtk_text(R, C, dx, depth, o.field("vv"), o.field("uu"), z)
        got = (outdir / f"Results_{tempo}.vtk").read_text()
        assert got == exp, f"Results_{tempo}.vtk differs"
        nfiles += 1
    assert nfiles == 5
    assert (outdir / "WaterLevel.txt").read_text() == "".join(wl)
    assert (outdir / f"TimeSimu_{tempo}.txt").exists()
    got = (pathlib.Path(base) / "EFDC" / "input" / "QSER.INP").read_text()
    assert got == qser_text(hist, CASE_KW["nlayers"], CASE_KW["basetime"])
    assert any(v > 0 for r in hist[1:] for e in ("north", "south", "east", "west") for v in r[e]["q"])
    # constituent series after the run: QSER's records, the north edge's times
    nrec = [len(hist[-1][e]["q"]) for e in ("north", "south", "east", "west")]
    times = [r["time"] if nrec[0] > 0 else 0.0 for r in hist]
    ev, nl, bt = CASE_KW["efdc_values"], CASE_KW["nlayers"], CASE_KW["basetime"]
    inp = pathlib.Path(base) / "EFDC" / "input"
    T = CASE_KW["evaporation"][1]
    assert (inp / "DSER.INP").read_text() == series_text("D", times, nrec, lambda t: ev["DYE_IN"], nl, bt)
    assert (inp / "SSER.INP").read_text() == series_text("S", times, nrec, lambda t: ev["SALT_IN"], nl, bt)
    assert (inp / "TSER.INP").read_text() == series_text("T", times, nrec, lambda t: T[t], nl, bt)
    assert (inp / "SHOW.INP").is_file() and (inp / "TEMP.INP").is_file()
    # EFDC.INP, rewritten with the first output: C23's record counts and
    # C24's records per edge are QSER.INP's series
    from test_efdc_config import EfdcReader
    efdc = (inp / "EFDC.INP").read_bytes()
    rd = EfdcReader(efdc)
    rd.seek("C23")
    c23 = [int(v) for v in rd.read(12)]
    assert c23[4] == c23[6] == sum(nrec) > 0
    qser = (inp / "QSER.INP").read_text()
    for e, n in zip(("North", "South", "East", "West"), nrec):
        assert qser.count(f"' *** {e}_edge_") == n
        assert efdc.decode("latin-1").count(f"! {e}_bc") == n
    rd.seek("C24")
    ij = [tuple(int(v) for v in rd.read(14)[:2]) for _ in range(sum(nrec))]
    assert ij == [(int(i), int(j)) for e in ("north", "south", "east", "west")
                  for i, j in zip(hist[0][e]["I"], hist[0][e]["J"])]


@pytest.mark.gpu
def test_driver_reports_an_output_write_failure(cases, tmp_path):
    """An output file that cannot be written (here a directory stands where
    Results_0.vtk goes) ends the run with status 1 and a message, after the
    writer thread has drained (no silent loss of outputs)."""
    spec = small_case(cases)
    base = cases.write_db(spec, tmp_path, "C5", **CASE_KW)
    outdir = pathlib.Path(base) / "SW2D" / "output"
    outdir.mkdir(parents=True, exist_ok=True)
    (outdir / "Results_0.vtk").mkdir()
    r = run_driver(tmp_path, check=False)
    assert r.returncode == 1
    assert "could not be written" in r.stdout
    assert (outdir / "Results_1.vtk").is_file()   # the other outputs were written


@pytest.mark.gpu
@pytest.mark.parametrize("after", [0, 2])
def test_driver_reports_a_step_failure(cases, tmp_path, after):
    """A library error during the run (injected: once `after` outputs are
    out, the next step request asks for a negative iteration count, which
    sw2d_step rejects) ends the run with status 1 and a message; the outputs
    taken before it are written, no later ones, and no TimeSimu file claims a
    finished run."""
    spec = small_case(cases)
    base = pathlib.Path(cases.write_db(spec, tmp_path, "C6", **CASE_KW))
    r = run_driver(tmp_path, "--fail-step-after-outputs", str(after), check=False)
    assert r.returncode == 1, r.stdout[-2000:]
    assert "step failed" in r.stdout and "the run stopped" in r.stdout
    outdir = base / "SW2D" / "output"
    assert len(list(outdir.glob("Results_*.vtk"))) == after
    assert not list(outdir.glob("TimeSimu_*.txt"))
    # the same case without the injection runs to the end
    ok = run_driver(tmp_path)
    assert ok.returncode == 0 and list(outdir.glob("TimeSimu_*.txt"))


def vtk_block(text, name):
    """Values of one SCALARS block of a STRUCTURED_POINTS file (file order)."""
    lines = text.splitlines()
    i = lines.index(f"SCALARS {name} float 1") + 2
    n = int(next(l for l in lines if l.startswith("POINT_DATA")).split()[1])
    return np.array([float(v) for v in lines[i:i + n]])


@pytest.mark.gpu
def test_driver_fp32_depths_within_a_millimetre(cases, po, tmp_path):
    """--precision fp32 (FP32 state, FP64 balance) through the whole drop-in
    run: every VTK depth within 1 mm of the oracle's FP64 depth at the same
    output (the files print 6 decimals)."""
    spec = small_case(cases)
    base = cases.write_db(spec, tmp_path, "C8", **CASE_KW)
    run_driver(tmp_path, "--precision", "fp32")
    R, C, dx = spec.z.shape[0], spec.z.shape[1], spec.cellsize
    z = po.lake_catchment_smooth(spec.lake_mask, spec.z)
    o = po.Oracle(z, spec.h0, dx, dt=spec.dt, manning=spec.manning, INT=0.1, INF=0.05, LWL=0.2,
                  EV_WL_min=0.01)
    o.set_rain(spec.rain, spec.dtrain)
    xs, ys, Q, dtoq = CASE_KW["sources"]
    o.set_sources([po.source_index(R, C, 0.0, 0.0, dx, xs[0], ys[0])], Q, dtoq)
    albedo, T, Rg = CASE_KW["evaporation"]
    o.set_evaporation(po.evaporation(albedo, T, Rg, spec.dtrain))
    dt, lkout = spec.dt, round(CASE_KW["dkout"] / spec.dt)
    time, mstep, out0, worst, n = 0.0, 0, -1, 0.0, 0
    outdir = pathlib.Path(base) / "SW2D" / "output"
    while time + dt <= spec.t_end:
        out0 += 1
        o.step(1)
        time += 2 * dt
        mstep += 2
        if not (mstep % lkout == 0 or out0 == 0) or out0 == 0:
            continue
        tempo = str(int(round(mstep * dt / CASE_KW["dkout"])))
        got = vtk_block((outdir / f"Results_{tempo}.vtk").read_text(), "depth")
        want = o.field("h")[::-1].ravel()
        worst = max(worst, float(np.max(np.abs(got - want))))
        n += 1
    assert n == 4 and worst <= 1e-3, (n, worst)


@pytest.mark.parametrize("ncells", [1, 10, 100])
def test_efdc_initial_fields_edge_widths(cases, tmp_path, ncells):
    """DYE/SALT/TEMP.INP numbering widths at the powers of ten (1 lake cell
    gives width 0, 10 and 100 cells the next widths) with one layer."""
    spec = small_case(cases)
    R, C = spec.z.shape
    mask = np.zeros((R, C), np.int32)
    idx = np.random.default_rng(ncells).choice(R * C, ncells, replace=False)
    mask.ravel()[idx] = 1
    spec.lake_mask = mask
    kw = dict(CASE_KW, nlayers=1)
    kw["efdc_values"] = dict(kw["efdc_values"], NWSER=0)
    base = pathlib.Path(cases.write_db(spec, tmp_path, "C9", **kw))
    run_driver(tmp_path, "--efdc-inputs")
    ev = kw["efdc_values"]
    got = (base / "EFDC" / "input" / "DYE.INP").read_text()
    assert got == initial_field_text(mask, 1, lambda r, i, l: ev["DYE_INIT"])
    assert len(got.splitlines()) == 5 + ncells
    tb, ts = ev["TEMP_BOT"], ev["TEMP_SURF"]
    temp = lambda r, i, l: tb + (ts - tb) / (1.0 + (spec.h0[r, i] / 1) * (0 - l))
    assert (base / "EFDC" / "input" / "TEMP.INP").read_text() == initial_field_text(mask, 1, temp)
    assert not (base / "EFDC" / "input" / "WSER.INP").exists()   # NWSER = 0
