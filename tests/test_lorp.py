"""EFDC's domain decomposition file LORP.INP (the reference runs its Gorp
partitioner on CELL.INP and renames LORP_sc_<N>_{h,v,r}.INP by setup.dat's
LORP type, SW2D-EFDC.cu:4480-4507; Gorp.h:461-739), written by
`sw2d-efdc --efdc-inputs` into EFDC/input/LORP.INP and by
`sw2d-efdc --lorp CELL.INP N TYPE OUT`.

Checked
- byte for byte against the files the reference's own partitioner wrote for
  the reference case's CELL.INP with its setup's 6 ranks (all three types;
  tests/golden/make_lorp_fixtures.py made them with oracle/_ref/gorp, the
  reference's Gorp sources compiled as its Makefile compiles them);
- when oracle/_ref/gorp is built (it needs the reference tree), against it on
  random lake masks of 3..160 x 3..260 cells, 1..10 ranks, every type;
- the cost model on a single tile (N = 1) against its closed form, and the
  file's structure (widths summing to the grid plus the ghost zones, active
  ids and the rank table consistent, tile counts by type) on every case.
"""
import os
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
DRIVER = pathlib.Path(os.environ.get("SW2D_DRIVER") or ROOT / "sw2d-efdc_amd" / "sw2d-efdc")
GORP = ROOT / "oracle" / "_ref" / "gorp"
GOLDEN = ROOT / "tests" / "golden"
TAGS = {1: "h", 2: "v", 3: "r"}


def write_cell(path, t):
    """A CELL.INP holding the types t (rows from the north) in the layout the
    partitioner reads: title, three header lines, 120-column blocks of rows
    with the digits from character 6."""
    R, C = t.shape
    with open(path, "w") as f:
        f.write(f"C cell.inp file, {C} columns and {R} rows @\nC\nC\nC\n")
        for b in range(0, C, 120):
            for i in range(R):
                f.write("%3d  " % (R - i) + "".join(str(v) for v in t[i, b:b + 120]) + "\n")
        f.write("C\n")


def random_types(rng):
    R, C = int(rng.integers(3, 160)), int(rng.integers(3, 260))
    n = rng.random((R, C))
    for _ in range(int(rng.integers(0, 4))):
        n = (n + np.roll(n, 1, 0) + np.roll(n, -1, 0) + np.roll(n, 1, 1) + np.roll(n, -1, 1)) / 5
    lake = n > np.quantile(n, rng.uniform(0.3, 0.95))
    edge = np.zeros_like(lake)
    edge[1:] |= lake[1:] != lake[:-1]
    edge[:-1] |= lake[:-1] != lake[1:]
    edge[:, 1:] |= lake[:, 1:] != lake[:, :-1]
    edge[:, :-1] |= lake[:, :-1] != lake[:, 1:]
    t = np.where(lake, 5, np.where(edge, 9, 0))
    if rng.random() < 0.2:
        t[rng.random((R, C)) < 0.05] = 9
    return t


def run_lorp(cell, n, kind, out):
    r = subprocess.run([str(DRIVER), "--lorp", str(cell), str(n), str(kind), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return pathlib.Path(out).read_text()


def parse(text):
    lines = text.split("\n")
    assert lines[-1] == ""
    head = lines[3].split()
    nh, nv, nact = int(head[0]), int(head[1]), int(head[2])
    k = lines.index("*   IC:    an array of NPARTX size setting x-width for each partition")
    ic = [int(v) for v in lines[k + 1:k + 1 + nh]]
    assert lines[k + 1 + nh] == "*   JC:    an array of NPARTY size setting y-height for each partition"
    jc = [int(v) for v in lines[k + 2 + nh:k + 2 + nh + nv]]
    k = k + 2 + nh + nv
    assert lines[k] == "*   List of active partition IDs, 0-based"
    ids = [int(v) for v in lines[k + 1:k + 1 + nact]]
    k = k + 1 + nact
    assert lines[k].startswith("*   Lookup table to go from your partition id to node id")
    table = [int(v) for v in lines[k + 1:-1]]
    return dict(nh=nh, nv=nv, nact=nact, cost=float(head[3]), weights=[float(v) for v in head[4:]],
                ic=ic, jc=jc, ids=ids, table=table, lines=lines)


def check_structure(p, R, C, n, kind):
    assert p["lines"][0] == f"LORP_sc_{n}_{TAGS[kind]}.INP"
    assert p["lines"][1] == (" LORP input data used for decomposing domain into optimal configuration "
                             f"with {n} processors")
    assert sum(p["ic"]) == C + 4 * p["nh"] and sum(p["jc"]) == R + 4 * p["nv"]
    if kind == 1:
        assert (p["nh"], p["nv"]) == (n, 1)
    elif kind == 2:
        assert (p["nh"], p["nv"]) == (1, n)
    else:
        assert p["nh"] * p["nv"] >= n
    assert len(p["ids"]) == p["nact"] and p["ids"] == sorted(set(p["ids"]))
    assert len(p["table"]) == p["nh"] * p["nv"]
    expect = [-1] * len(p["table"])
    for rank, i in enumerate(p["ids"]):
        expect[i] = rank
    assert p["table"] == expect
    assert p["weights"] == pytest.approx([0, 0, 0.001 / 276, 0.007 / 44 / 25, 0.003 / 44 / 25], rel=1e-5)


def reference_case_cell(cases, tmp_path, nproc=6, lorp=3):
    g = np.load(GOLDEN / "peri_lake.npz")
    mask = g["mask"]
    R, C = mask.shape
    spec = cases.config0(rows=R, cols=C, seed=5)
    spec.t_end = 1200.0
    spec.rain = np.zeros(4)
    spec.dtrain = 600.0
    spec.lake_mask = mask
    spec.h0 = g["h0"]
    base = pathlib.Path(cases.write_db(spec, tmp_path, "G1", efdc=True, basetime=3600.0,
                                       efdc_values={"nprocessors": nproc, "lorp": lorp},
                                       grid_make=(GOLDEN / "peri_lake_grid_make_input.dat").read_text()))
    r = subprocess.run([str(DRIVER), "--root", str(tmp_path), "--efdc-inputs"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return base, R, C


def test_lorp_of_the_reference_case(cases, tmp_path):
    """The driver's EFDC/input/LORP.INP for the reference setup (6 ranks, type
    r) and the other two types, equal to the reference partitioner's files."""
    base, R, C = reference_case_cell(cases, tmp_path)
    got = (base / "EFDC" / "input" / "LORP.INP").read_text()
    assert got == (GOLDEN / "peri_lake_LORP_sc_6_r.INP").read_text()
    check_structure(parse(got), R, C, 6, 3)
    cell = base / "EFDC" / "input" / "gefdc" / "CELL.INP"
    for kind in (1, 2):
        text = run_lorp(cell, 6, kind, tmp_path / f"L{kind}.INP")
        assert text == (GOLDEN / f"peri_lake_LORP_sc_6_{TAGS[kind]}.INP").read_text()
        check_structure(parse(text), R, C, 6, kind)


def test_single_tile_cost_closed_form(tmp_path):
    """N = 1, type h: one tile over the grid; its cost is the model's
    ghost-perimeter + lake (scaled to the ghost-ring area) + boundary terms."""
    rng = np.random.default_rng(3)
    for case in range(4):
        t = random_types(rng)
        R, C = t.shape
        write_cell(tmp_path / "CELL.INP", t)
        p = parse(run_lorp(tmp_path / "CELL.INP", 1, 1, tmp_path / "L.INP"))
        lake, bnd = int((t == 5).sum()), int((t == 9).sum())
        cost = (0.001 / 276 * (2 * R + 2 * C) + 0.007 / 44 / 25 * (lake / (R * C) * ((R + 4) * (C + 4)))
                + 0.003 / 44 / 25 * bnd) if lake + bnd else 0.0
        assert p["cost"] == pytest.approx(cost, rel=5e-6)
        assert p["ic"] == [C + 4] and p["jc"] == [R + 4]
        assert p["ids"] == ([0] if cost > 0 else []) and p["nact"] == len(p["ids"])


def test_bad_rank_count_is_reported(tmp_path):
    write_cell(tmp_path / "CELL.INP", np.full((5, 6), 5))
    r = subprocess.run([str(DRIVER), "--lorp", str(tmp_path / "CELL.INP"), "0", "3", str(tmp_path / "L")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "at least 1" in r.stdout


def test_malformed_cell_inp_is_reported(tmp_path):
    """A CELL.INP without the title line, or one that ends inside a block of
    rows, is an error with a message, not a partition of garbage."""
    (tmp_path / "A.INP").write_text("no title here\nC\nC\nC\n")
    t = np.full((6, 9), 5)
    write_cell(tmp_path / "B.INP", t)
    lines = (tmp_path / "B.INP").read_text().splitlines()
    (tmp_path / "B.INP").write_text("\n".join(lines[:7]) + "\n")   # 3 of 6 rows
    for name, msg in (("A.INP", "title"), ("B.INP", "ends inside a block")):
        r = subprocess.run([str(DRIVER), "--lorp", str(tmp_path / name), "2", "3", str(tmp_path / "L")],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stdout, (name, r.stdout)


@pytest.mark.skipif(not GORP.exists(), reason="oracle/_ref/gorp not built (needs the reference tree)")
def test_lorp_equals_the_reference_partitioner(tmp_path):
    rng = np.random.default_rng(11)
    for case in range(24):
        t = random_types(rng)
        R, C = t.shape
        n = int(rng.integers(1, 11))
        d = tmp_path / f"c{case}"
        d.mkdir()
        write_cell(d / "CELL.INP", t)
        subprocess.run([str(GORP), "CELL.INP", "sc", str(n)], cwd=d, capture_output=True, timeout=600,
                       check=True)
        for kind, tag in TAGS.items():
            ours = run_lorp(d / "CELL.INP", n, kind, d / f"ours_{tag}.INP")
            assert ours == (d / f"LORP_sc_{n}_{tag}.INP").read_text(), (case, R, C, n, tag)
            check_structure(parse(ours), R, C, n, kind)
