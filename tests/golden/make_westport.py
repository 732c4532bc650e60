"""Writes tests/golden/westport_100x100.npz: the reference's example inputs
(data.dat:35-45 -> example/westport_bathy_example.nc, NetCDF classic, and
example/new_westport_runupB_{h,u,v}.nc, NetCDF-4) loaded on the data.dat
grid of 100 x 100 cells by the oracle restatement of the Read_files = 1
loader, with the data.dat physical parameters.  The fixture is data only:
the initial state, fixed-bed flags, boundary series and times.

Run from the repository root where the reference tree is present:
    python tests/golden/make_westport.py [reference root]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import pyingest as OI  # noqa: E402

# data.dat:12-31 (the reference's example configuration)
DATA_DAT = dict(total_time=7200.0, save_time=20.0, npx=100, npy=100, CFL=0.9, epsilon_h=1e-4, R=2.6, DS=2e-4,
                FR=2.5e-2, porosity=0.4, OC=0.0047, Dsg=100e-6, sea_level_correction=0.0)


def main(ref="/root/reference"):
    ex = os.path.join(ref, "example")
    d = OI.load_problem_files(os.path.join(ex, "westport_bathy_example.nc"),
                              os.path.join(ex, "new_westport_runupB_h.nc"), "0",
                              DATA_DAT["sea_level_correction"], DATA_DAT["npx"], DATA_DAT["npy"])
    out = os.path.join(ROOT, "tests", "golden", "westport_100x100.npz")
    extent = np.array([d["xmin"], d["xmax"], d["ymin"], d["ymax"], d["dx"], d["dy"]])
    np.savez_compressed(out, h=d["h"], H=d["H"], fixH=d["fixH"], CC=d["CC"], tCC=d["tCC"], extent=extent,
                        data_dat=np.array([DATA_DAT[k] for k in sorted(DATA_DAT)]),
                        data_dat_keys=np.array(sorted(DATA_DAT)))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(*sys.argv[1:])
