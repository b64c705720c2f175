"""The grid generator's inputs of a coupled run (make_gefdc_inputs,
SW2D-EFDC.cu:2831-3680, called at 4472): CELL.INP, gridext.inp, depdat.inp
and gefdc.inp, written by `sw2d-efdc --efdc-inputs` into EFDC/input/gefdc/.

Checked against a statement of the rules independent of the writer's loops:
- cell types (whole-array numpy): contour flag = lake cell, or mask value
  different from a 4-neighbour in the grid; type 5 for lake cells, else 9
  when the flag is set or differs from a neighbour's (north/south in the
  grid, east/west inside the 120-column block the cell is printed in), else
  0;
- CELL.INP read back the way the Gorp decomposition tool reads it (title
  "C <name> file, <cols> columns and <rows> rows", three more header lines,
  then per block the rows from the north with the digits from character 5),
  which also pins the row-number column width for fewer than 1000 rows;
- depdat.inp: one line per lake cell, block by block and row by row, at the
  cell centre, with the bed after lake_catchment_smooth (the oracle's) as a
  float;
- gridext.inp: every cell in the same order, I, J and its south-west corner;
- gefdc.inp: the values of grid_make_input.dat and NDEPDAT = lake cells, read
  back by card.
On the reference case's mask (701 x 556: five blocks) with its own
grid_make_input.dat, and on small masks with lake cells on the grid border
and 120 / 121 / 241 columns (an empty last block, one-column blocks).  The
comment texts of CELL.INP's column-number lines and of gefdc.inp are parity
unpinned beyond these checks: the reference ships no output of this step.
"""
import os
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
DRIVER = pathlib.Path(os.environ.get("SW2D_DRIVER") or ROOT / "sw2d-efdc_amd" / "sw2d-efdc")
GOLDEN = ROOT / "tests" / "golden"


def expected_types(mask):
    m = np.asarray(mask)
    R, C = m.shape
    diff = np.zeros((R, C), bool)
    diff[1:] |= m[1:] != m[:-1]
    diff[:-1] |= m[:-1] != m[1:]
    diff[:, 1:] |= m[:, 1:] != m[:, :-1]
    diff[:, :-1] |= m[:, :-1] != m[:, 1:]
    flag = (m == 1) | diff
    near = np.zeros((R, C), bool)
    near[1:] |= flag[1:] != flag[:-1]
    near[:-1] |= flag[:-1] != flag[1:]
    j = np.arange(C)
    b = (j // 120) * 120
    e = np.minimum(b + 120, C)
    west = (j > 0) & ((j != b) | (j == e - 1))
    east = j < e - 1
    wd = np.zeros((R, C), bool)
    wd[:, 1:] = flag[:, 1:] != flag[:, :-1]
    ed = np.zeros((R, C), bool)
    ed[:, :-1] = flag[:, :-1] != flag[:, 1:]
    near |= (wd & west[None, :]) | (ed & east[None, :])
    t = np.where(flag | near, 9, 0)
    return np.where(m == 1, 5, t)


def read_cell_like_gorp(path):
    lines = open(path).read().split("\n")
    head = lines[0].split()
    assert head[0] == "C" and head[2] == "file," and head[4] == "columns" and head[5] == "and"
    cols, rows = int(head[3]), int(head[6])
    t = np.full((rows, cols), -1)
    k, wtot = 4, 0
    while wtot < cols:
        wcur = None
        for jr in range(rows):
            buf = lines[k]
            k += 1
            digits = ""
            for ch in buf[5:]:
                if not ch.isdigit():
                    break
                digits += ch
            if wcur is None:
                wcur = len(digits)
            t[jr, wtot:wtot + len(digits)] = [int(c) for c in digits]
        assert wcur > 0
        wtot += wcur
    return t, lines


def run_inputs(cases, tmp_path, spec, **kw):
    base = pathlib.Path(cases.write_db(spec, tmp_path, "G1", efdc=True, basetime=3600.0, **kw))
    r = subprocess.run([str(DRIVER), "--root", str(tmp_path), "--efdc-inputs"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return base / "EFDC" / "input" / "gefdc"


def grid_make_values(text):
    vals = {}
    for line in text.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v)
    return vals


def check(gdir, mask, z_smooth, res, gm):
    R, C = mask.shape
    exp = expected_types(mask)
    got, lines = read_cell_like_gorp(gdir / "CELL.INP")
    assert np.array_equal(got, exp)
    assert lines[0] == f"C cell.inp file, {C} columns and {R} rows @"
    assert lines[3] == "C"
    # block order: block by block, rows from the north, columns west to east
    order = [(i, j) for b in range(0, C, 120) for i in range(R) for j in range(b, min(b + 120, C))]
    g = np.loadtxt(gdir / "gridext.inp", ndmin=2)
    assert g.shape == (R * C, 4)
    ii = np.array([o[0] for o in order])
    jj = np.array([o[1] for o in order])
    assert np.array_equal(g[:, 0], jj + 1) and np.array_equal(g[:, 1], R - ii)
    np.testing.assert_allclose(g[:, 2], jj * res, atol=5e-7)
    np.testing.assert_allclose(g[:, 3], (R - ii - 1) * res, atol=5e-7)
    lake = [(i, j) for (i, j) in order if mask[i, j] == 1]
    d = np.loadtxt(gdir / "depdat.inp", ndmin=2)
    assert d.shape == (len(lake), 3) and len(lake) == int((mask == 1).sum())
    li = np.array([p[0] for p in lake])
    lj = np.array([p[1] for p in lake])
    np.testing.assert_allclose(d[:, 0], (lj + 0.5) * res, atol=5e-7)
    np.testing.assert_allclose(d[:, 1], (R - li - 0.5) * res, atol=5e-7)
    zf = z_smooth[li, lj].astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(d[:, 2], zf, atol=5e-7)
    # gefdc.inp: the value line after each card's header line
    gl = (gdir / "gefdc.inp").read_text().splitlines()

    def card(name):
        i = [k for k, l in enumerate(gl) if l.startswith(name + " ")][1]
        return [float(v) for v in gl[i + 1].split()]

    assert card("C2") == [0, 0, 1, C, 1, R, C, R]
    assert card("C3") == pytest.approx([0, 0, 0, res, res, 1])
    assert card("C5") == pytest.approx([gm[k] for k in ("ITRXM", "ITRHM", "ITRKM", "ITRGM", "NDEPSM",
                                                         "NDEPSMF", "DEPMIN", "DDATADJ")])
    assert card("C6") == pytest.approx([gm[k] for k in ("RPX", "RPK", "RPH", "RSQXM", "RSQKM", "RSQKIM",
                                                         "RSQHM", "RSQHIM", "RSQHJM")], rel=1e-6)
    assert card("C7") == pytest.approx([gm[k] for k in ("XSHIFT", "YSHIFT", "HSCALE", "RKJDKI",
                                                         "ANGORO")])
    assert card("C8") == [gm[k] for k in ("ISIRKI", "JSIRKI", "ISIHIHJ", "JSIHIHJ")]
    c11 = card("C11")
    assert c11[1] == len(lake)
    assert c11 == pytest.approx([gm["ISIDEP"], len(lake), gm["CDEP"], gm["RADM"], gm["ISIDPTYP"],
                                 gm["SURFELV"], gm["ISVEG"], gm["NVEGDAT"], gm["NVEGTYP"],
                                 gm["ZROUGH"], gm["VERBOSE"]], rel=1e-6)


def test_gefdc_inputs_on_the_reference_case(cases, po, tmp_path):
    g = np.load(GOLDEN / "peri_lake.npz")
    mask = g["mask"]
    R, C = mask.shape
    spec = cases.config0(rows=R, cols=C, seed=5)
    spec.t_end = 1200.0
    spec.rain = np.zeros(4)
    spec.dtrain = 600.0
    spec.lake_mask = mask
    spec.h0 = g["h0"]
    gm_text = (GOLDEN / "peri_lake_grid_make_input.dat").read_text()
    gdir = run_inputs(cases, tmp_path, spec, grid_make=gm_text)
    z = po.lake_catchment_smooth(mask, spec.z)
    check(gdir, mask, z, spec.cellsize, grid_make_values(gm_text))


@pytest.mark.parametrize("cols", [17, 120, 121, 241])
def test_gefdc_inputs_block_edges(cases, po, tmp_path, cols):
    """Lake cells on every grid border and block edge; 120 columns leave an
    empty last block (its row numbers run together on one line, as the
    reference prints them), 121 and 241 a one-column last block."""
    R = 13
    rng = np.random.default_rng(cols)
    mask = (rng.random((R, cols)) < 0.35).astype(np.int32)
    mask[rng.random((R, cols)) < 0.1] = 2          # other mask classes
    mask[0, 3:5] = 1
    mask[-1, 2] = 1
    mask[4:7, 0] = 1
    mask[5, cols - 1] = 1
    mask[0, 0] = mask[0, -1] = mask[-1, 0] = mask[-1, -1] = 0   # corners: see the module text
    spec = cases.config0(rows=R, cols=cols, seed=9)
    spec.t_end = 1200.0
    spec.rain = np.zeros(4)
    spec.dtrain = 600.0
    spec.lake_mask = mask
    gdir = run_inputs(cases, tmp_path, spec)
    z = po.lake_catchment_smooth(mask, spec.z)
    if cols % 120 == 0:
        # the empty last block: row numbers only, on one line, which the
        # closing "C" line then continues
        text = (gdir / "CELL.INP").read_text().splitlines()
        assert "".join("%3d  " % (R - i) for i in range(R)) + "C" in text
    check(gdir, mask, z, spec.cellsize, grid_make_values(cases.GRID_MAKE_INPUT))


def test_cell_inp_header_under_ten_columns(cases, tmp_path):
    """Fewer than ten columns: the reference's second column-number line has
    no digits and no line end, so it runs into the "C" line after it."""
    spec = cases.config0(rows=12, cols=7, seed=9)
    spec.t_end = 1200.0
    spec.rain = np.zeros(4)
    spec.dtrain = 600.0
    m = np.zeros((12, 7), np.int32)
    m[4:6, 2:5] = 1
    spec.lake_mask = m
    gdir = run_inputs(cases, tmp_path, spec)
    text = (gdir / "CELL.INP").read_text().splitlines()
    assert text[:3] == ["C cell.inp file, 7 columns and 12 rows @", "C    0        1", "C    C"]
    rows = [l[5:] for l in text[3:15]]
    assert [len(r) for r in rows] == [7] * 12
    assert np.array_equal(np.array([[int(c) for c in r] for r in rows]), expected_types(m))


def test_missing_grid_make_input_is_reported(cases, tmp_path):
    spec = cases.config0(rows=10, cols=12, seed=9)
    spec.t_end = 1200.0
    spec.rain = np.zeros(4)
    spec.dtrain = 600.0
    m = np.zeros((10, 12), np.int32)
    m[4:6, 4:8] = 1
    spec.lake_mask = m
    base = pathlib.Path(cases.write_db(spec, tmp_path, "G2", efdc=True, basetime=3600.0))
    (base / "grid_make_input.dat").unlink()
    r = subprocess.run([str(DRIVER), "--root", str(tmp_path), "--efdc-inputs"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "grid_make_input.dat" in r.stdout
