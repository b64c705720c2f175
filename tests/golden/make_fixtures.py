"""Write the test fixtures derived from the reference's example case data.

Reads db/COUPLED_peri_lake/SW2D/input/{mask.dat, initial_condition.dat} of
LabHidro/SW2D-EFDC (a checkout at /root/reference by default) and stores them
as compressed arrays in tests/golden/peri_lake.npz.  These are input data
files of the reference's only example case (556 x 701, dx = 10 m); its DEM is
not part of the repository, so no flow results exist to compare against.

The case's EFDC card template (EFDC/input/template/EFDC.INP, the file the
coupled run copies and rewrites cards of) and its EFDC time-series locations
(SW2D/input/efdc_coord_out.dat) are stored as they are, the template
gzip-compressed: peri_lake_EFDC.INP.gz, peri_lake_efdc_coord_out.dat; so is
its grid generator's control file (grid_make_input.dat):
peri_lake_grid_make_input.dat.
"""
import gzip
import pathlib
import shutil
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def read_esri(path, dtype):
    with open(path) as f:
        hdr = {}
        for _ in range(6):
            k, v = f.readline().split()
            hdr[k.lower()] = float(v)
        vals = np.array(f.read().split(), dtype=np.float64)
    return hdr, vals.astype(dtype)


def main(ref="/root/reference"):
    case = pathlib.Path(ref) / "db" / "COUPLED_peri_lake" / "SW2D" / "input"
    hdr, mask = read_esri(case / "mask.dat", np.int32)
    R, C = int(hdr["nrows"]), int(hdr["ncols"])
    mask = mask[: R * C].reshape(R, C)
    h0 = np.array(open(case / "initial_condition.dat").read().split(), dtype=np.float64)
    h0 = h0[: R * C].reshape(R, C)
    np.savez_compressed(HERE / "peri_lake.npz", mask=mask, h0=h0,
                        header=np.array([hdr["ncols"], hdr["nrows"], hdr["xllcorner"],
                                         hdr["yllcorner"], hdr["cellsize"], hdr["nodata_value"]]))
    print(R, C, int((mask == 1).sum()), h0.max())
    root = pathlib.Path(ref) / "db" / "COUPLED_peri_lake"
    raw = (root / "EFDC" / "input" / "template" / "EFDC.INP").read_bytes()
    with gzip.GzipFile(HERE / "peri_lake_EFDC.INP.gz", "wb", mtime=0) as g:
        g.write(raw)
    shutil.copyfile(case / "efdc_coord_out.dat", HERE / "peri_lake_efdc_coord_out.dat")
    shutil.copyfile(root / "grid_make_input.dat", HERE / "peri_lake_grid_make_input.dat")
    print(len(raw), "bytes of EFDC.INP")


if __name__ == "__main__":
    main(*sys.argv[1:])
