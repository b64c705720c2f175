"""EFDC.INP card update of the coupled run (write_efdc_CONFIG,
SW2D-EFDC.cu:1069-2024, called with the first output at 4924-4933 on the
copy of EFDC/input/template/EFDC.INP made at 4281-4289).

The driver's --efdc-inputs writes the file without a GPU: the first output's
coupling records depend on mask.dat alone.  The file is read back the way
EFDC reads it: SEEK (EFDC_src/SEEK.f) finds the second line that starts with
the card's name followed by a blank (the title, then the header), and a
list-directed READ takes the values from the lines after it
(EFDC_src/SCANEFDC.f reads C6, C7, C9, C14, C22 and C23 so).  Checked:
  - every value EFDC reads from the seventeen rewritten cards, against the
    case (setup.dat's EFDC block, efdc_coord_out.dat, the mask);
  - C23's record counts and C24's cell list against the oracle's
    couple_BC_efdc records, edges N, S, E, W;
  - every line outside the rewritten cards is the template's line.
Templates: the reference case's own (tests/golden/peri_lake_EFDC.INP.gz, CRLF
lines) with its mask and time-series points, and a small synthetic one with
CRLF and with LF lines.  Beyond these checks the bytes of the rewritten cards
(comment texts, column spacing) are parity unpinned: the reference's
in-place rewrite leaves the end of the file damaged (it drops or adds a byte
per card), so no output of it is a usable golden file.
"""
import gzip
import os
import pathlib
import re
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
# SW2D_DRIVER selects another build of the driver (the sanitizer build)
DRIVER = pathlib.Path(os.environ.get("SW2D_DRIVER") or ROOT / "sw2d-efdc_amd" / "sw2d-efdc")
GOLDEN = ROOT / "tests" / "golden"

# rewritten cards and the card after each (the end of its block)
CARDS = [("C6", "C7"), ("C7", "C8"), ("C8", "C9"), ("C9", "C10"), ("C10", "C11"),
         ("C11", "C11A"), ("C14", "C15"), ("C22", "C23"), ("C23", "C24"), ("C24", "C25"),
         ("C25", "C26"), ("C26", "C27"), ("C71", "C71A"), ("C72", "C73"), ("C73", "C74"),
         ("C84", "C85"), ("C87", "C88")]

EV = dict(dye_on=1, temp_on=1, salt_on=0, NWSER=1, NASER=1, ZBRADJ=0.25, ZBRCVRT=1.5, HMIN=0.125,
          HADJ=-0.5, HCVRT=1.0, HDRY=0.02, HWET=0.04, BELADJ=2.5, BELCVRT=1.25, ISPPH=1,
          NPPPH=12, ISRPPH=0, IPPHXY=3, ISVPH=1, NPVPH=6, ISRVPH=1, IVPHXY=2, NTSSSS=1, MTSP=1,
          MTSC=1, MTSA=0, MTSUE=1, MTSUT=1, MTSU=0, NTSPTC=720)


class EfdcReader:
    """EFDC's input reading: SEEK and list-directed READ over the lines."""

    def __init__(self, data: bytes):
        # Fortran reads records split at LF; a CR stays in the record
        self.lines = data.decode("latin-1").split("\n")
        self.pos = 0

    def seek(self, tag):
        L = len(tag)
        hits = 0
        while self.pos < len(self.lines):
            text = self.lines[self.pos][:80]
            self.pos += 1
            t = text.lstrip(" ")
            if len(t.rstrip(" ")) < L:
                continue
            t = t.upper()
            if t[:L] == tag and t[L:L + 1] in (" ", ""):
                hits += 1
                if hits == 2:
                    return
        raise AssertionError(f"GROUP: {tag} NOT FOUND BEFORE END OF FILE")

    def read(self, n):
        """READ(1,*) of n items: values from as many records as needed."""
        items = []
        while len(items) < n:
            assert self.pos < len(self.lines), "end of file in a READ"
            rec = self.lines[self.pos].replace("\r", " ")
            self.pos += 1
            items += [v for v in re.split(r"[,\s]+", rec.strip()) if v]
        return items[:n]

    def ints(self, n):
        return [int(v) for v in self.read(n)]

    def reals(self, n):
        return [float(v) for v in self.read(n)]


def synthetic_template(cases, newline):
    t = cases.efdc_card_template()
    return t if newline == "\r\n" else t.replace(b"\r\n", b"\n")


def run_inputs(cases, tmp_path, spec, template, points, **kw):
    base = pathlib.Path(cases.write_db(
        spec, tmp_path, "E1", efdc=True, basetime=3600.0, nlayers=kw.pop("nlayers", 3),
        efdc_values=EV, efdc_points=points, tctmsr=60.0, efdc_template=template,
        dkout=600.0, dpout=600.0, **kw))
    r = subprocess.run([str(DRIVER), "--root", str(tmp_path), "--efdc-inputs"],
                       capture_output=True, text=True, timeout=600)
    return base, r


def card_blocks(lines):
    """Line ranges [title line of card, title line of its successor)."""
    first = {}
    for i, l in enumerate(lines):
        tok = l.split()
        if tok and tok[0] not in first:
            first[tok[0]] = i
    return [(first[a], first[b]) for a, b in CARDS]


def outside(lines):
    keep = np.ones(len(lines), bool)
    for a, b in card_blocks(lines):
        keep[a:b] = False
    out = [l.rstrip() for l, k in zip(lines, keep) if k]
    # the reference starts a card after the previous token and one
    # character, and some cards with a newline: a line of the template may
    # end early or be followed by an empty line there
    return [l for l in out if l]


def check_cards(data, template, *, R, C, nlayers, water, recs, lts, spec, timmax, tctmsr,
                ntsptc, out_outlet, tbegin):
    ev = EV
    rd = EfdcReader(data)
    rd.seek("C6")
    rows = [rd.ints(10) for _ in range(9)]
    salt, temp, dye = ev["salt_on"], ev["temp_on"], ev["dye_on"]
    assert rows[0] == [0] * 10
    assert rows[1][:2] == [salt, salt] and rows[2][:2] == [temp, 1] and rows[3][:2] == [dye, dye]
    assert all(r[2:5] == [0, 1, 1] and r[5:] == [0] * 5 for r in rows[1:])
    assert all(r[:2] == [0, 0] for r in rows[4:])
    ntc = int(timmax / 3600.0)
    rd.seek("C7")
    assert rd.ints(11) == [ntc, ntsptc, 0, 0, ntc, 0, 0, 0, ntc, 1, 16]
    rd.seek("C8")
    assert rd.reals(9) == [3600.0, tbegin, 3600.0, 0.0, 0, 0, 0, 0, 0.0]
    rd.seek("C9")
    v = rd.read(14)
    assert [int(x) for x in v[:12]] == [nlayers, C, R, water + 2, water, 1, 1, water, 0, 0, 0, 0]
    assert [float(x) for x in v[12:]] == [0.03125, 0.0625]
    rd.seek("C10")
    dz = [rd.reals(2) for _ in range(nlayers)]
    assert [int(k) for k, _ in dz] == list(range(1, nlayers + 1))
    assert all(abs(d - 1.0 / nlayers) <= 5e-7 for _, d in dz)
    rd.seek("C11")
    v = rd.reals(13)
    assert v == [spec.cellsize, spec.cellsize, 1.0, 0, ev["ZBRADJ"], ev["ZBRCVRT"], ev["HMIN"],
                 ev["HADJ"], ev["HCVRT"], ev["HDRY"], ev["HWET"], ev["BELADJ"], ev["BELCVRT"]]
    rd.seek("C14")
    assert rd.ints(10) == [0, ev["NWSER"], ev["NASER"], 0, 0, 0, 0, 0, 0, 0]
    nq = sum(len(r["I"]) for r in recs.values())
    rd.seek("C22")
    assert rd.ints(11) == [0, 0, 0, salt * nq, temp * nq, dye * nq, 0, 0, 0, 0, 0]
    rd.seek("C23")
    assert rd.ints(12) == [0, 0, 0, 0, nq, 0, nq, 0, 0, 0, 0, 0]
    rd.seek("C24")
    got = [rd.read(14) for _ in range(nq)]
    exp_ij = [(int(i), int(j)) for e in ("north", "south", "east", "west")
              for i, j in zip(recs[e]["I"], recs[e]["J"])]
    assert [(int(g[0]), int(g[1])) for g in got] == exp_ij
    for k, g in enumerate(got, start=1):
        assert float(g[2]) == 0.0 and [int(x) for x in g[3:5]] == [0, 0]
        assert [int(x) for x in g[5:13]] == [k, salt * k, temp * k, dye * k, 0, 0, 0, 0]
        assert float(g[13]) == 1.0
    # the edge comment of each record line
    c24 = EfdcReader(data)
    c24.seek("C24")
    tags = [c24.lines[c24.pos + i].split("!")[-1].strip() for i in range(nq)]
    exp_tags = [f"{e.capitalize()}_bc" for e in ("north", "south", "east", "west")
                for _ in recs[e]["I"]]
    assert tags == exp_tags
    rd.seek("C25")
    assert all(rd.reals(4) == [0.0] * 4 for _ in range(nq))
    rd.seek("C26")
    assert all(rd.ints(2) == [0, 0] for _ in range(nq))
    rd.seek("C71")
    c71 = [rd.ints(4) for _ in range(7)]
    assert [r[0] for r in c71] == [salt, temp, dye, 0, 0, 0, 0]
    assert [r[3] for r in c71] == [1, 1, 1, 3, 3, 3, 3] and all(r[1:3] == [1, 0] for r in c71)
    rd.seek("C72")
    assert rd.ints(4) == [ev["ISPPH"], ev["NPPPH"], ev["ISRPPH"], ev["IPPHXY"]]
    rd.seek("C73")
    assert rd.ints(4) == [ev["ISVPH"], ev["NPVPH"], ev["ISRVPH"], ev["IVPHXY"]]
    rd.seek("C84")
    v = rd.read(7)
    assert [int(x) for x in v[:6]] == [out_outlet, len(lts), 1, ntc * ntsptc, ntsptc, 1]
    assert float(v[6]) == tctmsr
    rd.seek("C87")
    for k, (i, j) in enumerate(lts):
        v = rd.read(12)
        assert [int(x) for x in v[:11]] == [i, j, ev["NTSSSS"], ev["MTSP"], ev["MTSC"], ev["MTSA"],
                                           ev["MTSUE"], ev["MTSUT"], ev["MTSU"], 0, 0]
        assert v[11] == f"'P_{k}'"
    # outside the rewritten cards the file is the template
    out_lines = data.decode("latin-1").split("\n")
    tpl_lines = template.decode("latin-1").split("\n")
    assert outside(out_lines) == outside(tpl_lines)


def small_spec(cases, R=30, C=41):
    spec = cases.config0(rows=R, cols=C, seed=3)
    spec.t_end = 7200.0
    spec.rain = np.zeros(int(spec.t_end // 600.0) + 2)
    spec.dtrain = 600.0
    mask = np.zeros((R, C), np.int32)
    mask[8:14, 10:25] = 1
    mask[14:20, 12:16] = 1
    mask[3, 0:3] = 1          # a lake touching the west edge
    spec.lake_mask = mask
    return spec


@pytest.mark.parametrize("newline", ["\r\n", "\n"])
def test_efdc_cards_on_a_synthetic_template(cases, po, tmp_path, newline):
    spec = small_spec(cases)
    template = synthetic_template(cases, newline)
    points = [(15.0, 25.0), (105.0, 255.0), (398.0, 2.0)]
    base, r = run_inputs(cases, tmp_path, spec, template, points)
    assert r.returncode == 0, r.stdout + r.stderr
    data = (base / "EFDC" / "input" / "EFDC.INP").read_bytes()
    R, C = spec.z.shape
    recs = po.couple_bc(spec.lake_mask, spec.h0, np.zeros((R, C)), np.zeros((R, C)),
                        spec.cellsize, spec.cellsize)
    assert sum(len(v["I"]) for v in recs.values()) > 20
    lts = [(int(np.floor(abs(0.0 - x) / spec.cellsize + 0.5)),
            int(np.floor(abs(0.0 - y) / spec.cellsize + 0.5))) for x, y in points]
    check_cards(data, template, R=R, C=C, nlayers=3, water=int((spec.lake_mask == 1).sum()),
                recs=recs, lts=lts, spec=spec, timmax=spec.t_end, tctmsr=60.0, ntsptc=720,
                out_outlet=0, tbegin=float(cases.efdc_value("TBEGIN", EV)))
    # the template under EFDC/input/template is left as it was
    assert (base / "EFDC" / "input" / "template" / "EFDC.INP").read_bytes() == template


def test_efdc_cards_on_the_reference_case_template(cases, po, tmp_path):
    """The reference case's template, mask (701 x 556, 50638 lake cells) and
    time-series points, on a flat bed."""
    g = np.load(GOLDEN / "peri_lake.npz")
    mask = g["mask"]
    R, C = mask.shape
    spec = cases.config0(rows=R, cols=C, seed=3)
    spec.cellsize = float(g["header"][4])
    spec.t_end = 7200.0
    spec.rain = np.zeros(int(spec.t_end // 600.0) + 2)
    spec.dtrain = 600.0
    spec.lake_mask = mask
    spec.h0 = g["h0"]
    template = gzip.decompress((GOLDEN / "peri_lake_EFDC.INP.gz").read_bytes())
    txt = (GOLDEN / "peri_lake_efdc_coord_out.dat").read_text().split()
    n = int(txt[3])
    points = [(float(txt[4 + 2 * i]), float(txt[5 + 2 * i])) for i in range(n)]
    base, r = run_inputs(cases, tmp_path, spec, template, points, nlayers=1)
    assert r.returncode == 0, r.stdout + r.stderr
    data = (base / "EFDC" / "input" / "EFDC.INP").read_bytes()
    recs = po.couple_bc(mask, spec.h0, np.zeros((R, C)), np.zeros((R, C)), spec.cellsize,
                        spec.cellsize)
    lts = [(int(np.floor(abs(x) / spec.cellsize + 0.5)), int(np.floor(abs(y) / spec.cellsize + 0.5)))
           for x, y in points]
    assert lts == [(492, 534), (493, 535)]
    check_cards(data, template, R=R, C=C, nlayers=1, water=50638, recs=recs, lts=lts, spec=spec,
                timmax=spec.t_end, tctmsr=60.0, ntsptc=720, out_outlet=0,
                tbegin=float(cases.efdc_value("TBEGIN", EV)))


def test_template_without_a_card_is_reported(cases, tmp_path):
    spec = small_spec(cases)
    template = cases.efdc_card_template().replace(b"C88 ", b"X88 ")
    base, r = run_inputs(cases, tmp_path, spec, template, [(15.0, 25.0)])
    assert r.returncode == 1 and "EFDC.INP has no card C88" in r.stdout


def test_missing_template_is_reported(cases, tmp_path):
    spec = small_spec(cases)
    base, _ = run_inputs(cases, tmp_path, spec, None, [(15.0, 25.0)])
    (base / "EFDC" / "input" / "template" / "EFDC.INP").unlink()
    r = subprocess.run([str(DRIVER), "--root", str(tmp_path), "--efdc-inputs"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 1 and "unknown file" in r.stdout and "template/EFDC.INP" in r.stdout
