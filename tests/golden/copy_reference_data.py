"""Copies the reference's own data files used as fixtures by the format
tests (run in the build container, where the reference tree is present):
the flat-plate example's control deck, grid (GRID1D, Fortran unformatted)
and boundary-layer inflow / upper data, and the jet and isotropic decks.
These are input data of the reference's examples, not code."""
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "flat225_opencfd-comb2.in": "examples/flat-2.25/opencfd-comb2.in",
    "flat225_OCFD-grid.dat": "examples/flat-2.25/OCFD-grid.dat",
    "flat225_flow1d-inlet-comb.dat": "examples/flat-2.25/flow1d-inlet-comb.dat",
    "flat225_flow0d-upper-comb.dat": "examples/flat-2.25/flow0d-upper-comb.dat",
    "jet_opencfd-comb2.in": "cases/jet/opencfd-comb2.in",
    "isotropic_opencfd-comb2.in": "examples/isotropic/opencfd-comb2.in",
}

if __name__ == "__main__":
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
        print(dst)
