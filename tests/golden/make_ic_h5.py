"""Writes tests/golden/bump_2d_ic.h5: an initial-condition file for tests/golden/bump_2d.msh
in the layout of the reference's case tools (make_bump_case.py:127-160: a flat, cell-major
float64 dataset "state" of ncells * 6 values in the mesh file's cell order, attributes
num_cells and num_components), written by tests/cpp/h5_ic_write.cpp through libhdf5's C API.

The values are closed forms of the cell centroids (ic_values below), so the tests recompute
them.  Run from the repository root: python tests/golden/make_ic_h5.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))

MESH = os.path.join(ROOT, "tests", "golden", "bump_2d.msh")
OUT = os.path.join(ROOT, "tests", "golden", "bump_2d_ic.h5")
NCOMP = 6


def ic_values(xy):
    """(ncells, 6) from centroids: bed bump, dam-break depth, and four distinct smooth fields."""
    x, y = xy[:, 0], xy[:, 1]
    xm = x.max()
    X = np.zeros((len(x), NCOMP))
    X[:, 0] = 0.1 * np.exp(-((x - 0.5 * xm) ** 2) / 0.05)
    X[:, 1] = np.where(x < 0.4 * xm, 1.0, 0.6) - X[:, 0]
    X[:, 2] = 0.01 * (1.0 + x)
    X[:, 3] = -0.005 * y
    X[:, 4] = 0.002 * x * y
    X[:, 5] = 0.001 + 0.0 * x
    return X


def input_order_centroids(mesh):
    xy = np.empty((mesh.ncells, 2))
    xy[mesh.cell_perm] = mesh.cell_centroid[:, :2]
    return xy


def main():
    from conftest import zdm
    mesh = zdm.Mesh.read_msh(MESH)
    vals = ic_values(input_order_centroids(mesh))
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "h5_ic_write")
        subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "h5_ic_write.cpp"),
                        "-ldl", "-o", exe], check=True)
        text = "\n".join(repr(float(v)) for v in vals.ravel())
        subprocess.run([exe, OUT, str(mesh.ncells), str(NCOMP)], input=text, text=True, check=True)
    print("wrote", OUT, mesh.ncells, "cells x", NCOMP)


if __name__ == "__main__":
    main()
