"""The C++ host CLI (the reference's solver executable over the MI355X library) against the
Python host on the same settings file: same state bitwise, since both drive the same kernels."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT, cases, zdm

pytestmark = pytest.mark.gpu
MESH = os.path.join(ROOT, "tests", "golden", "bump_2d.msh")


def input_order_state(mesh, Xs):
    X = np.empty_like(Xs)
    X[mesh.cell_perm] = Xs
    return X


def run_cli(model, settings, tmp_path, *extra):
    sp = tmp_path / "settings.json"
    sp.write_text(json.dumps(settings))
    out = tmp_path / "out.npy"
    cmd = [os.path.join(PKG_DIR, "zdm_cli_%s" % model), "-settings", str(sp), "-o", str(out)] + list(extra)
    print("case cli", model, " ".join(extra), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    return np.load(out), r.stdout


@pytest.mark.parametrize("positivity,ti", [("none", "splitting"), ("mood", "splitting"), ("zhang_shu", "splitting"),
                                           ("none", "imex"), ("mood", "imex"), ("none", "implicit")])
def test_cli_muscl_equals_python_solver(tmp_path, positivity, ti):
    mesh = zdm.Mesh.read_msh(MESH)
    xy = mesh.cell_centroid[:, :2]
    Xs = np.zeros((mesh.ncells, 6))
    Xs[:, 0] = 0.1 * np.exp(-((xy[:, 0] - 0.5 * xy[:, 0].max()) ** 2) / 0.05)
    Xs[:, 1] = np.where(xy[:, 0] < 0.4 * xy[:, 0].max(), 1.0, 0.6) - Xs[:, 0]
    Xs[:, 2] = 0.01
    np.save(tmp_path / "ic.npy", input_order_state(mesh, Xs))
    settings = {"name": "cli", "io": {"mesh_path": MESH, "initial_condition_file": str(tmp_path / "ic.npy")},
                "solver": {"t_end": 1e9, "cfl": 0.45, "reconstruction_order": 2, "ssp_stages": 2,
                           "limiter": "tvd", "positivity": positivity, "time_integration": ti}}
    Xc, log = run_cli("sme1", settings, tmp_path, "-max_steps", "25", "-no_output")
    assert "Finished" in log
    s = zdm.Solver("sme1", mesh, settings=zdm.Settings(settings))
    s.upload(Xs)
    s.initialize()
    s.step(25)
    Xp, _ = s.download()
    s.close()
    assert np.array_equal(Xc, input_order_state(mesh, Xp))


@pytest.mark.parametrize("mask", [None, [0, 1]])
def test_cli_hdf5_initial_condition_equals_python_solver(tmp_path, mask):
    """io.initial_condition_file = an HDF5 "state" file (tests/golden/bump_2d_ic.h5, written by
    tests/cpp/h5_ic_write.cpp) with and without initial_condition_mask: 20 steps of the CLI and
    of the Python solver started from zdm.initial_state on the same settings, bitwise equal."""
    io = {"mesh_path": MESH, "initial_condition_file": os.path.join(ROOT, "tests", "golden", "bump_2d_ic.h5")}
    if mask is not None:
        io["initial_condition_mask"] = mask
    settings = {"name": "cli", "io": io,
                "solver": {"t_end": 1e9, "cfl": 0.45, "reconstruction_order": 2, "ssp_stages": 2, "limiter": "tvd"}}
    Xc, log = run_cli("sme1", settings, tmp_path, "-max_steps", "20", "-no_output")
    assert "Loading initial condition" in log
    mesh = zdm.Mesh.read_msh(MESH)
    X0 = zdm.initial_state(settings, mesh, "sme1")
    s = zdm.Solver("sme1", mesh, settings=zdm.Settings(settings))
    s.upload(X0)
    s.initialize()
    s.step(20)
    Xp, _ = s.download()
    s.close()
    assert np.all(np.isfinite(Xp))
    assert np.array_equal(Xc, input_order_state(mesh, Xp))


def test_cli_chorin_equals_python_solver(tmp_path):
    mesh = zdm.Mesh.read_msh(MESH)
    Xs = cases.c3_state(mesh)
    np.save(tmp_path / "ic.npy", input_order_state(mesh, Xs))
    settings = {"name": "cli", "io": {"mesh_path": MESH, "initial_condition_file": str(tmp_path / "ic.npy")},
                "solver": {"t_end": 1e9, "cfl": 0.3, "reconstruction_order": 1, "method": "chorin"}}
    Xc, log = run_cli("vam", settings, tmp_path, "-max_steps", "3", "-ksp_gmres_restart", "30",
                      "-ksp_rtol", "1e-10", "-pc_type", "pbjacobi")
    s = zdm.Solver("vam", mesh, settings=zdm.Settings(settings))
    s.chorin_setup(restart=30, max_it=400, rtol=1e-10, atol=1e-12)
    s.upload(Xs)
    s.initialize()
    n, its, reason = s.chorin_step(3)
    Xp, _ = s.download()
    s.close()
    assert n == 3
    assert np.array_equal(Xc, input_order_state(mesh, Xp))


def test_cli_refuses_what_it_does_not_provide(tmp_path):
    settings = {"io": {"mesh_path": MESH}, "solver": {"method": "chorin", "t_end": 1e9}}
    sp = tmp_path / "s.json"
    sp.write_text(json.dumps(settings))
    r = subprocess.run([os.path.join(PKG_DIR, "zdm_cli_vam"), "-settings", str(sp), "-pc_type", "gamg"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    settings["solver"]["method"] = "muscl"
    settings["solver"]["time_integration"] = "crank_nicolson"
    sp.write_text(json.dumps(settings))
    r = subprocess.run([os.path.join(PKG_DIR, "zdm_cli_sme1"), "-settings", str(sp)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "crank_nicolson" in r.stderr


def read_vtu(path):
    """VTK XML UnstructuredGrid with appended raw data (UInt64 headers): {array name: data}."""
    import re
    raw = open(path, "rb").read()
    head, _, data = raw.partition(b'<AppendedData encoding="raw">')
    data = data[data.index(b"_") + 1:]
    out = {}
    for m in re.finditer(rb'<DataArray type="(\w+)"(?: Name="(\w+)")?(?: NumberOfComponents="(\d+)")?'
                         rb' format="appended" offset="(\d+)"', head):
        typ, name, ncomp, off = m.group(1).decode(), (m.group(2) or b"Points").decode(), int(m.group(3) or 1), int(m.group(4))
        nbytes = int(np.frombuffer(data[off:off + 8], dtype=np.uint64)[0])
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[typ]
        arr = np.frombuffer(data[off + 8:off + 8 + nbytes], dtype=dt)
        out[name] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return out


@pytest.mark.parametrize("model,method", [("sme1", "muscl"), ("vam", "chorin")])
def test_cli_vtk_snapshots(tmp_path, model, method):
    """io.snapshots VTK output (IOManager.hpp:100-120, 346-376): '<filename>-NNN.vtu' every
    t_end / snapshots plus the .vtu.series file; the last snapshot holds the final state and
    the mesh of the input file in its cell order."""
    mesh = zdm.Mesh.read_msh(MESH)
    xy = mesh.cell_centroid[:, :2]
    info = zdm.ModelInfo(model)
    Xs = cases.c3_state(mesh) if model == "vam" else np.zeros((mesh.ncells, info.n_q))
    if model != "vam":
        Xs[:, 1] = np.where(xy[:, 0] < 0.4 * xy[:, 0].max(), 1.0, 0.6)
    np.save(tmp_path / "ic.npy", input_order_state(mesh, Xs))
    settings = {"name": "vtk", "io": {"mesh_path": MESH, "initial_condition_file": str(tmp_path / "ic.npy"),
                                      "directory": str(tmp_path / "out"), "filename": "run", "snapshots": 4},
                "solver": {"t_end": 0.02, "cfl": 0.3, "reconstruction_order": 1, "method": method}}
    extra = ["-ksp_gmres_restart", "30", "-ksp_rtol", "1e-8"] if method == "chorin" else []
    Xc, log = run_cli(model, settings, tmp_path, *extra)
    series = json.loads((tmp_path / "out" / "run.vtu.series").read_text())
    files = series["files"]
    times = [f["time"] for f in files]
    assert files[0]["name"] == "run-000.vtu" and times[0] == 0.0
    assert len(files) == 5 and all(b > a for a, b in zip(times, times[1:]))
    for k, t in enumerate(times[1:], 1):
        assert t >= k * 0.02 / 4 - 1e-9 * 0.005
    v = read_vtu(tmp_path / "out" / files[-1]["name"])
    assert v["State"].shape == (mesh.ncells, info.n_q) and v["Aux"].shape[0] == mesh.ncells
    assert np.array_equal(v["State"], Xc)
    v0 = read_vtu(tmp_path / "out" / "run-000.vtu")
    assert v0["Points"].shape[1] == 3 and v0["connectivity"].max() < v0["Points"].shape[0]
    nvc = np.diff(np.concatenate([[0], v0["offsets"]]))
    assert len(nvc) == mesh.ncells and np.all(v0["types"] == np.where(nvc == 3, 5, 9))
    assert np.array_equal(v0["State"], input_order_state(mesh, Xs))
