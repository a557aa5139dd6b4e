"""Convert the reference's advection_2d HDF5 regression fixtures into compact
golden vectors (level, i, j, u) for the tests in this directory.

Source files (BSD-3-Clause, hpc-maths/samurai v0.33.0):
  tests/reference/finite_volume/
    test_finite_volume_advection_2d_finite-volume-advection-2d-0.01_pred_{0,1}{,_init}.h5
They are written by demos/FiniteVolume/advection_2d.cpp run with ``--Tf 0.01``
(tests/test_demo_finite_volume.py:55-72) through io/hdf5.hpp:84-145: one quad per
leaf cell in for_each_cell(mesh[cells]) order (levels ascending, rows, x).

A second fixture, advection_2d_save.json, keeps what pins the HDF5/XDMF
writer (io/hdf5.hpp) against these files: per file the numbers of cells and
points, the SHA-256 of /mesh/points (little-endian float64 [n_points][3]) and of
/mesh/connectivity (little-endian uint64 [n_cells][4]), and the text of the
.xdmf file written beside it.

Decoding: the quad's first point is the cell's min corner (x0, y0); its edge
length is 2**-level on the unit square, so level = -log2(len), i = x0/len,
j = y0/len, all exact in binary.  Only the container's h5dump tool is used to
read the files (no reference code is run).

Usage: python tests/golden/make_advection_2d_fixtures.py [reference_root]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

H5DUMP = "/opt/conda/bin/h5dump"
SRC_DIR = "tests/reference/finite_volume"
STEM = "test_finite_volume_advection_2d_finite-volume-advection-2d-0.01_pred_{s}{suffix}.h5"


def _dump(h5file, dset, dtype, tmp):
    out = os.path.join(tmp, "d.bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", out, h5file],
                   check=True, stdout=subprocess.DEVNULL)
    return np.fromfile(out, dtype=dtype)


def convert(h5file):
    with tempfile.TemporaryDirectory() as tmp:
        u = _dump(h5file, "/mesh/fields/u", "<f8", tmp)
        pts = _dump(h5file, "/mesh/points", "<f8", tmp).reshape(-1, 3)
        con = _dump(h5file, "/mesh/connectivity", "<u8", tmp).reshape(-1, 4)
    assert con.shape[0] == u.shape[0]
    p = pts[con.astype(np.int64)]            # (n, 4, 3)
    x0 = p[:, :, 0].min(axis=1)
    y0 = p[:, :, 1].min(axis=1)
    length = p[:, :, 0].max(axis=1) - x0
    level = np.rint(-np.log2(length)).astype(np.int8)
    assert np.all(np.ldexp(1.0, -level.astype(np.int64)) == length)
    i = np.rint(x0 / length).astype(np.int32)
    j = np.rint(y0 / length).astype(np.int32)
    assert np.all(i * length == x0) and np.all(j * length == y0)
    return level, i, j, u


def save_record(h5file):
    with tempfile.TemporaryDirectory() as tmp:
        pts = _dump(h5file, "/mesh/points", "<f8", tmp)
        con = _dump(h5file, "/mesh/connectivity", "<u8", tmp)
    with open(h5file[:-3] + ".xdmf") as f:
        xdmf = f.read()
    return {"n_cells": con.size // 4, "n_points": pts.size // 3,
            "points_sha256": hashlib.sha256(pts.astype("<f8").tobytes()).hexdigest(),
            "connectivity_sha256": hashlib.sha256(con.astype("<u8").tobytes()).hexdigest(),
            "xdmf": xdmf}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    records = {}
    for s in (0, 1):
        for suffix in ("_init", ""):
            src = os.path.join(ref, SRC_DIR, STEM.format(s=s, suffix=suffix))
            level, i, j, u = convert(src)
            name = f"advection_2d_pred{s}{'_init' if suffix else '_t0.01'}.npz"
            np.savez_compressed(os.path.join(here, name), level=level, i=i, j=j, u=u)
            print(name, level.size, "cells, levels", level.min(), "..", level.max())
            records[os.path.basename(src)[:-3]] = save_record(src)
    with open(os.path.join(here, "advection_2d_save.json"), "w") as f:
        json.dump(records, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
