"""Write the LORP.INP fixtures of the reference case: the files the
reference's own domain partitioner (Gorp, compiled from the reference tree
into oracle/_ref/gorp by `make -C oracle ref`) writes for the CELL.INP that
`sw2d-efdc --efdc-inputs` makes of the reference case's mask, with the
reference setup's 6 ranks: peri_lake_LORP_sc_6_{h,v,r}.INP.

Run from the repository root after building the driver:
  make -C oracle ref && python tests/golden/make_lorp_fixtures.py
"""
import importlib.util
import pathlib
import shutil
import subprocess
import sys
import tempfile

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def main():
    d = ROOT / "sw2d-efdc_amd"
    spec = importlib.util.spec_from_file_location("sw2d_efdc_amd", d / "__init__.py",
                                                  submodule_search_locations=[str(d)])
    pkg = importlib.util.module_from_spec(spec)
    sys.modules["sw2d_efdc_amd"] = pkg
    spec.loader.exec_module(pkg)
    from sw2d_efdc_amd import cases
    sys.path.insert(0, str(ROOT / "tests"))
    from test_lorp import reference_case_cell
    with tempfile.TemporaryDirectory() as tmp:
        tmp = pathlib.Path(tmp)
        base, R, C = reference_case_cell(cases, tmp)
        cell = base / "EFDC" / "input" / "gefdc" / "CELL.INP"
        run = tmp / "gorp"
        run.mkdir()
        shutil.copyfile(cell, run / "CELL.INP")
        subprocess.run([str(ROOT / "oracle" / "_ref" / "gorp"), "CELL.INP", "sc", "6"], cwd=run, check=True)
        for tag in "hvr":
            shutil.copyfile(run / f"LORP_sc_6_{tag}.INP", HERE / f"peri_lake_LORP_sc_6_{tag}.INP")
            print(tag, (HERE / f"peri_lake_LORP_sc_6_{tag}.INP").read_text().splitlines()[3])


if __name__ == "__main__":
    main()
