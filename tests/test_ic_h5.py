"""HDF5 initial conditions (IOManager::LoadSolution, IOManager.hpp:122-192; selected by
io.initial_condition_file / io.initial_condition_mask, VirtualSolver.hpp:234-281), CPU only.

The fixture tests/golden/bump_2d_ic.h5 was written by tests/cpp/h5_ic_write.cpp (libhdf5's
C API, independent of the library's reader) in the layout of the reference's case tools:
a flat cell-major float64 dataset "state" in the mesh file's cell order.  Its values are
closed forms of the cell centroids (tests/golden/make_ic_h5.py), so the expected state in
the solver's (renumbered) cell order is recomputed here; the Python reader and the C++ host
(CLI -ic_only) must both reproduce it exactly, with and without a component mask."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT, zdm

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_ic_h5  # noqa: E402  (the fixture's closed forms)

GOLD = os.path.join(ROOT, "tests", "golden")
MESH = os.path.join(GOLD, "bump_2d.msh")
IC = os.path.join(GOLD, "bump_2d_ic.h5")
CLI = os.path.join(PKG_DIR, "zdm_cli_sme1")


@pytest.fixture(scope="module")
def mesh():
    return zdm.Mesh.read_msh(MESH)


def expected_solver_order(mesh):
    return make_ic_h5.ic_values(mesh.cell_centroid[:, :2])


def settings(ic, mask=None):
    io = {"mesh_path": MESH, "initial_condition_file": ic}
    if mask is not None:
        io["initial_condition_mask"] = mask
    return {"name": "ic", "io": io, "solver": {"t_end": 1.0}}


def test_fixture_reads_as_written(mesh):
    d = zdm.read_h5(IC)
    assert d.shape == (mesh.ncells * 6,)
    assert np.array_equal(d.reshape(-1, 6), make_ic_h5.ic_values(make_ic_h5.input_order_centroids(mesh)))


@pytest.mark.parametrize("mask", [None, [], [0, 1], [1, 4, 9]])
def test_python_initial_state(mesh, mask):
    X = zdm.initial_state(settings(IC, mask), mesh, "sme1")
    want = expected_solver_order(mesh)
    sel = list(range(6)) if not mask else [k for k in mask if k < 6]
    assert np.array_equal(X[:, sel], want[:, sel])
    rest = [k for k in range(6) if k not in sel]
    assert np.all(X[:, rest] == 0)


def run_cli_ic(tmp_path, s):
    sp = tmp_path / "s.json"
    sp.write_text(json.dumps(s))
    out = tmp_path / "ic.npy"
    r = subprocess.run([CLI, "-settings", str(sp), "-ic_only", "-o", str(out)], capture_output=True, text=True,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=""), timeout=120)
    return r, (np.load(out) if r.returncode == 0 else None)


@pytest.mark.skipif(not os.path.exists(CLI), reason="zdm_cli_sme1 is not built")
@pytest.mark.parametrize("mask", [None, [0, 1]])
def test_cli_reads_the_same_state(tmp_path, mesh, mask):
    r, Xin = run_cli_ic(tmp_path, settings(IC, mask))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Loading initial condition" in r.stdout
    Xs = zdm.initial_state(settings(IC, mask), mesh, "sme1")
    Xp = np.empty_like(Xs)
    Xp[mesh.cell_perm] = Xs                     # the CLI writes the mesh file's cell order
    assert np.array_equal(Xin, Xp)


def test_two_dimensional_dataset_and_round_trip(tmp_path, mesh):
    want = expected_solver_order(mesh)
    Xin = np.empty_like(want)
    Xin[mesh.cell_perm] = want
    p = str(tmp_path / "ic2.h5")
    zdm.write_h5(p, Xin)                        # (ncells, n_q), as PETSc's viewer writes a Vec of block size n_q
    assert np.array_equal(zdm.read_h5(p), Xin)
    assert np.array_equal(zdm.initial_state(settings(p), mesh, "sme1"), want)


def test_errors(tmp_path, mesh):
    short = str(tmp_path / "short.h5")
    zdm.write_h5(short, np.zeros(10))
    with pytest.raises(zdm.ZdmError, match="need 960 x 6"):
        zdm.initial_state(settings(short), mesh, "sme1")
    other = str(tmp_path / "other.h5")
    zdm.write_h5(other, np.zeros(5760), dataset="solution")
    with pytest.raises(zdm.ZdmError, match='no dataset "state"'):
        zdm.initial_state(settings(other), mesh, "sme1")
    with pytest.raises(zdm.ZdmError, match="cannot be opened"):
        zdm.initial_state(settings(str(tmp_path / "missing.h5")), mesh, "sme1")
    if os.path.exists(CLI):
        for path, msg in ((short, "need 960 x 6"), (other, 'no dataset "state"'),
                          (str(tmp_path / "missing.h5"), "cannot be opened")):
            r, _ = run_cli_ic(tmp_path, settings(path))
            assert r.returncode != 0 and msg in r.stderr, r.stderr


def test_missing_library_is_reported(tmp_path):
    """ZDM_HDF5_LIB naming no library: the reader fails with a message naming it (a fresh
    process, since the library is looked up once per process); the CLI reports it too."""
    code = ("import sys; sys.path.insert(0, %r); from conftest import zdm\n"
            "try:\n    zdm.read_h5(%r)\nexcept zdm.ZdmError as e:\n    print('ERR', e)\n") % (
        os.path.join(ROOT, "tests"), IC)
    env = dict(os.environ, ZDM_HDF5_LIB="/nonexistent/libhdf5.so", ZDM_NO_TORCH="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert "ERR HDF5 initial conditions need libhdf5" in r.stdout and "/nonexistent/libhdf5.so" in r.stdout, \
        r.stdout + r.stderr
    if os.path.exists(CLI):
        sp = tmp_path / "s.json"
        sp.write_text(json.dumps(settings(IC)))
        r = subprocess.run([CLI, "-settings", str(sp), "-ic_only"], capture_output=True, text=True,
                           env=dict(env, HIP_VISIBLE_DEVICES=""), timeout=120)
        assert r.returncode != 0 and "need libhdf5" in r.stderr
