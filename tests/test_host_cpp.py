"""The C++ host (include/zoomy/, zoomy-dmplex_amd/host/main.cpp) without a GPU: settings
parsing against the reference's keys and defaults (Settings.hpp:13-138) on the reference's
own settings files, and the error paths of the CLI."""
import json
import os
import subprocess

import pytest

from conftest import PKG_DIR, ROOT, zdm

GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(PKG_DIR, "zdm_cli_sme1")


def cli_settings(path):
    r = subprocess.run([CLI, "-settings", path, "-print-settings"], capture_output=True, text=True,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.fixture(scope="module", autouse=True)
def need_cli():
    if not os.path.exists(CLI):
        pytest.skip("zdm_cli_sme1 is not built")


@pytest.mark.parametrize("name", ["settings.json", "settings_bump_2d.json", "cfg_o2_n60.json"])
def test_reference_settings_files_parse_like_the_python_reader(name):
    path = os.path.join(GOLD, name)
    c = cli_settings(path)
    p = zdm.Settings.from_json(path)
    raw = json.load(open(path))
    assert c["name"] == raw["name"]
    for k in ("t_end", "cfl", "reconstruction_order", "min_dt", "limiter", "time_integration",
              "method", "positivity", "ssp_stages"):
        assert c["solver"][k] == pytest.approx(p.solver[k]) if isinstance(p.solver[k], float) \
            else c["solver"][k] == p.solver[k], k
    assert c["io"]["mesh_path"] == raw["io"]["mesh_path"]
    assert c["io"]["initial_condition_mask"] == raw["io"]["initial_condition_mask"]
    assert c["io"]["snapshots"] == raw["io"]["snapshots"]


def test_defaults_of_an_empty_settings_file(tmp_path):
    f = tmp_path / "s.json"
    f.write_text("{}")
    c = cli_settings(str(f))
    # Settings.hpp:13-40 defaults
    assert c["name"] == "Simulation"
    s = c["solver"]
    assert (s["t_end"], s["cfl"], s["reconstruction_order"], s["min_dt"]) == (1.0, 0.5, 1, 1e-12)
    assert (s["limiter"], s["time_integration"], s["method"], s["positivity"], s["ssp_stages"]) == \
        ("venkatakrishnan", "splitting", "muscl", "none", 5)
    assert c["io"]["directory"] == "output" and c["io"]["snapshots"] == 10


def test_json_reader_details(tmp_path):
    f = tmp_path / "s.json"
    f.write_text('{"name": "a\\"b\\u00e9\\n", "solver": {"t_end": 20, "reconstruction_order": 2.0,'
                 ' "cfl": 3e-1, "unknown": [1, {"x": null}], "method": "chorin"},'
                 ' "io": {"initial_condition_mask": [0, 1.5, "x", 3]},'
                 ' "model": {"parameters": {"g": 9.0, "nu": "no", "lam": 2}}}')
    c = cli_settings(str(f))
    assert c["name"] == 'a"bé\n'
    assert c["solver"]["t_end"] == 20.0 and c["solver"]["reconstruction_order"] == 2
    assert c["solver"]["cfl"] == 0.3 and c["solver"]["method"] == "chorin"
    assert c["io"]["initial_condition_mask"] == [0, 3]          # integers only
    assert c["parameters"] == {"g": 9.0, "lam": 2.0}             # numbers only


def test_cli_errors(tmp_path):
    r = subprocess.run([CLI, "-settings", str(tmp_path / "missing.json"), "-print-settings"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Could not open settings file" in r.stderr
    f = tmp_path / "bad.json"
    f.write_text('{"solver": {"t_end": }')
    r = subprocess.run([CLI, "-settings", str(f), "-print-settings"], capture_output=True, text=True)
    assert r.returncode == 1 and "json" in r.stderr
    f.write_text('{"solver": {"t_end": "late"}}')
    r = subprocess.run([CLI, "-settings", str(f), "-print-settings"], capture_output=True, text=True)
    assert r.returncode == 1 and "number" in r.stderr


def test_vtu_snapshots_written_by_the_io_manager(tmp_path):
    """zoomy::IOManager (IOManager.hpp:19-120, 346-376) on the CPU: snapshots every t_end /
    snapshots from time 0, '<filename>-NNN.vtu' with the mesh in its file's cell order and the
    cell data, and the .vtu.series listing."""
    import shutil
    import numpy as np
    from test_gpu_host_cpp import MESH, read_vtu
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "vtu_write")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "vtu_write.cpp"), "-L", PKG_DIR, "-lzdm_mesh",
                    "-Wl,-rpath," + PKG_DIR, "-o", exe], check=True, capture_output=True, text=True)
    r = subprocess.run([exe, MESH, str(tmp_path / "o")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    series = json.loads((tmp_path / "o" / "snap.vtu.series").read_text())
    assert [f["name"] for f in series["files"]] == ["snap-000.vtu", "snap-001.vtu", "snap-002.vtu"]
    assert [f["time"] for f in series["files"]] == [0.0, 0.5, 1.0]
    v = read_vtu(tmp_path / "o" / "snap-001.vtu")
    mesh = zdm.Mesh.read_msh(MESH)
    n = mesh.ncells
    assert np.array_equal(v["State"], 10.0 * np.arange(n)[:, None] + np.arange(3)[None, :])
    assert np.array_equal(v["Aux"], -np.arange(n, dtype=np.float64))
    nvc = np.diff(np.concatenate([[0], v["offsets"]]))
    assert len(nvc) == n and np.all(v["types"] == np.where(nvc == 3, 5, 9))
    # every cell's centroid (file order) lies in the bounding box of its vertices
    P = v["Points"].reshape(-1, 3)
    cen = np.empty((n, 3))
    cen[mesh.cell_perm] = mesh.cell_centroid
    for i in range(n):
        vs = P[v["connectivity"][v["offsets"][i] - nvc[i]:v["offsets"][i]]]
        assert np.all(vs.min(axis=0)[:2] - 1e-12 <= cen[i, :2]) and np.all(cen[i, :2] <= vs.max(axis=0)[:2] + 1e-12)
