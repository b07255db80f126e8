"""C ABI: every entry point declared in include/*.h is exported by the built libraries (no
compute calls).  Settings: the reference's JSON keys and defaults."""
import ctypes
import json
import os
import re

import pytest

from conftest import ROOT, zdm

MODELS = ["sme0_1d", "sme0", "sme0_visc", "sme1", "sme1m", "sme2", "vam"]


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(zdm_[a-z0-9_]+)\s*\(", txt)))


def test_mesh_library_exports_header():
    lib = ctypes.CDLL(os.path.join(ROOT, "zoomy-dmplex_amd", "libzdm_mesh.so"))
    names = declared("zdm_mesh.h")
    assert "zdm_mesh_tri_grid" in names
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("model", MODELS)
def test_model_library_exports_header(model):
    lib = ctypes.CDLL(os.path.join(ROOT, "zoomy-dmplex_amd", "libzdm_%s.so" % model))
    names = declared("zdm.h")
    assert "zdm_step" in names and "zdm_transport_rhs" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(zdm.SOLVER_SYMBOLS) <= set(names)


@pytest.mark.parametrize("model,nq,naux,dim", [("sme0_1d", 3, 4, 1), ("sme0", 4, 1, 2),
                                                ("sme1", 6, 1, 2), ("sme2", 8, 1, 2)])
def test_model_info_host_only(model, nq, naux, dim):
    info = zdm.ModelInfo(model)
    assert (info.n_q, info.n_aux, info.dim) == (nq, naux, dim)
    assert info.boundary_tags == ["default"]
    if model == "sme0_1d":
        # Model.H:43-44
        assert info.parameter_names == ["g", "rho", "nu", "lambda_s", "e_x"]
        assert list(info.default_params) == [9.81, 1.0, 1.0, 0.5, 0.0]


def test_settings_defaults_and_overrides(tmp_path):
    s = zdm.Settings()
    # Settings.hpp:13-40 defaults
    assert s.solver["t_end"] == 1.0 and s.solver["cfl"] == 0.5
    assert s.solver["reconstruction_order"] == 1 and s.solver["ssp_stages"] == 5
    assert s.solver["limiter"] == "venkatakrishnan" and s.solver["min_dt"] == 1e-12
    assert s.solver["time_integration"] == "splitting" and s.solver["method"] == "muscl"
    assert s.io["mesh_label"] == "Face Sets" and s.io["snapshots"] == 10
    d = {"name": "x", "solver": {"cfl": 0.45, "ssp_stages": 2, "dt": 0.1, "bogus": 1},
         "io": {"initial_condition_mask": [0, 1, "a"]},
         "model": {"parameters": {"g": 10.0, "lambda": 0.2, "zz": 1, "flag": "s"}}}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(d))
    s = zdm.Settings.from_json(str(p))
    assert s.solver["cfl"] == 0.45 and s.solver["ssp_stages"] == 2
    assert "dt" not in s.solver and "bogus" not in s.solver     # unknown keys ignored
    assert s.io["initial_condition_mask"] == [0, 1]
    assert s.parameters == {"g": 10.0, "lambda": 0.2, "zz": 1.0}
    with pytest.raises(zdm.ZdmError):
        zdm.Settings.from_json(str(tmp_path / "missing.json"))


def test_missing_library_fails_loudly():
    with pytest.raises(zdm.ZdmError):
        zdm.model_lib("no_such_model")
