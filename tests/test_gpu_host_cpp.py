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
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.load(out), r.stdout


def test_cli_muscl_equals_python_solver(tmp_path):
    mesh = zdm.Mesh.read_msh(MESH)
    xy = mesh.cell_centroid[:, :2]
    Xs = np.zeros((mesh.ncells, 6))
    Xs[:, 0] = 0.1 * np.exp(-((xy[:, 0] - 0.5 * xy[:, 0].max()) ** 2) / 0.05)
    Xs[:, 1] = np.where(xy[:, 0] < 0.4 * xy[:, 0].max(), 1.0, 0.6) - Xs[:, 0]
    Xs[:, 2] = 0.01
    np.save(tmp_path / "ic.npy", input_order_state(mesh, Xs))
    settings = {"name": "cli", "io": {"mesh_path": MESH, "initial_condition_file": str(tmp_path / "ic.npy")},
                "solver": {"t_end": 1e9, "cfl": 0.45, "reconstruction_order": 2, "ssp_stages": 2,
                           "limiter": "tvd"}}
    Xc, log = run_cli("sme1", settings, tmp_path, "-max_steps", "25")
    assert "Finished" in log
    s = zdm.Solver("sme1", mesh, settings=zdm.Settings(settings))
    s.upload(Xs)
    s.initialize()
    s.step(25)
    Xp, _ = s.download()
    s.close()
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
    settings["solver"]["time_integration"] = "imex"
    sp.write_text(json.dumps(settings))
    r = subprocess.run([os.path.join(PKG_DIR, "zdm_cli_sme1"), "-settings", str(sp)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "imex" in r.stderr
