"""Write tests/golden/pcc_conductionQ.npz from the reference's own fixture
files for the 1-D conduction model (tests/data/thermal/pcc_thermal_model_1d/
Tprofile_conductionQ and Tsurface_conductionQ, read by the reference's
tests/test_thermal.py:20-34).  They hold the PCC model's output, rounded to
3 decimals, for the case of tests/test_thermal.py:36-80 (60 layers,
zfac 1.05, zmax 2.5, ti 120, rhoc 960000, emissivity 1, Fgeotherm 0.2,
albedo 0.2 at 1.52 AU and 5 deg latitude, dt = 495661.779 s, 50001 steps).

Stored: the last temperature profile (61 values) and every 8th row of the
surface series [time / period, T[0], T[nz]] (one row per 3 steps), with the
step index of each kept row.  Run from the repository root:
    python tests/golden/make_pcc_fixture.py /root/reference
"""
import csv
import os
import sys

import numpy as np


def read_rows(path):
    rows = []
    with open(path) as fh:
        for row in csv.reader(fh, delimiter=" ", quoting=csv.QUOTE_NONNUMERIC):
            rows.append([x for x in row if x != ''])
    return np.array(rows, dtype=np.float64)


def main(ref):
    d = os.path.join(ref, "tests", "data", "thermal", "pcc_thermal_model_1d")
    prof = read_rows(os.path.join(d, "Tprofile_conductionQ"))
    surf = read_rows(os.path.join(d, "Tsurface_conductionQ"))
    keep = np.arange(0, surf.shape[0], 8)
    if keep[-1] != surf.shape[0] - 1:
        keep = np.append(keep, surf.shape[0] - 1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcc_conductionQ.npz")
    np.savez_compressed(out, Tprofile_last=prof[-1], Tsurface=surf[keep],
                        Tsurface_step=(3 * keep).astype(np.int64))
    print(out, prof.shape, surf.shape, keep.size)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
