"""The plug-in surface of the generated headers (CPU).

- The reference's own generated Model.H and Numerics.H compile unchanged into the product's
  kernel source (zoomy-dmplex_amd/csrc/zdm_solver.hip, hipcc for gfx950, host and device) and
  into the oracle, against the product's include/zoomy/UserFunctions.H.  Syntax only: nothing
  of the reference is built into an object, linked or run.  Skipped where the reference
  checkout is absent.
- A model without update_variables compiles (the clamp is optional, TransportStep.hpp:23-28).
- RefreshDerivativeAux (TransportStep.hpp:161-202) through the 1-D model's field-level
  update_aux_variables (Model.H:250-262): on a uniform line the Green-Gauss derivative with a
  zero-Neumann boundary has the closed form (q[i+1] - q[i-1]) / (2 dx) inside and
  (q[1] - q[0]) / (2 dx), (q[n-1] - q[n-2]) / (2 dx) at the two ends; with the setting off the
  derivative slots pass through.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT, oracle, zdm

REFERENCE = "/root/reference"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _model_dir_from(tmp_path, sources):
    d = tmp_path / "model"
    d.mkdir()
    for name, src in sources.items():
        if isinstance(src, str) and os.path.exists(src):
            os.symlink(src, d / name)
        else:
            (d / name).write_text(src)
    return str(d)


def _hipcc_syntax(model_dir):
    cmd = [HIPCC, "--offload-arch=gfx950", "-std=c++17", "-fsyntax-only", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
           "-Wno-unused-value", "-I" + os.path.join(ROOT, "include"), "-I" + model_dir,
           "-I" + os.path.join(PKG_DIR, "csrc"), "-I" + os.path.join(PKG_DIR, "models"),
           '-DZDM_MODEL_NAME="check"', os.path.join(PKG_DIR, "csrc", "zdm_solver.hip")]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600)


def _gxx_oracle_syntax(model_dir):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           "-I" + model_dir, "-I" + os.path.join(PKG_DIR, "models"), "-I" + os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "oracle", "oracle.cpp")]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600)


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "Model.H")), reason="no reference checkout")
@pytest.mark.skipif(not os.path.exists(HIPCC), reason="no hipcc")
def test_reference_generated_headers_compile_into_the_kernels(tmp_path):
    d = _model_dir_from(tmp_path, {"Model.H": os.path.join(REFERENCE, "Model.H"),
                                   "Numerics.H": os.path.join(REFERENCE, "Numerics.H")})
    r = _hipcc_syntax(d)
    assert r.returncode == 0, r.stderr[-3000:]
    r = _gxx_oracle_syntax(d)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="no hipcc")
def test_model_without_update_variables_compiles(tmp_path):
    src = open(os.path.join(PKG_DIR, "models", "sme0_1d", "Model.H")).read()
    a = src.index("    PORTABLE_FN static inline SimpleArray<T, 3> update_variables(")
    b = src.index("    PORTABLE_FN static inline SimpleArray<T, 4> update_aux_variables(")
    d = _model_dir_from(tmp_path, {"Model.H": src[:a] + src[b:],
                                   "Numerics.H": os.path.join(PKG_DIR, "models", "sme0_1d", "Numerics.H")})
    r = _hipcc_syntax(d)
    assert r.returncode == 0, r.stderr[-3000:]
    r = _gxx_oracle_syntax(d)
    assert r.returncode == 0, r.stderr[-3000:]


def _line_case(n=40, seed=4):
    rng = np.random.default_rng(seed)
    mesh = zdm.Mesh.line(n, 0.0, 2.0, renumber=False)
    X = np.zeros((n, 3))
    X[:, 0] = 0.1 * rng.uniform(size=n)
    X[:, 1] = 1.0 + 0.2 * rng.uniform(size=n)
    X[:, 2] = 0.3 * rng.standard_normal(n)
    A = rng.uniform(size=(n, 4))
    return mesh, X, A


def _gg_closed_form(q, dx):
    d = np.empty_like(q)
    d[1:-1] = (q[2:] - q[:-2]) / (2 * dx)
    d[0] = (q[1] - q[0]) / (2 * dx)
    d[-1] = (q[-1] - q[-2]) / (2 * dx)
    return d


def test_refresh_derivative_aux_closed_form_1d():
    mesh, X, A = _line_case()
    x = mesh.cell_centroid[:, 0]
    assert np.all(np.diff(x) > 0)
    dx = x[1] - x[0]
    orc = oracle("sme0_1d")
    Xc, Ac = X.copy(), A.copy()
    orc.enforce(mesh, orc.default_params, Xc, Ac)
    # the derivative aux is taken from the state before its clamp (TransportStep.hpp:202)
    for slot, field in ((0, 2), (1, 1), (2, 0)):
        ref = _gg_closed_form(X[:, field], dx)
        assert np.allclose(Ac[:, slot], ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref))), slot
    # KP inverse depth in slot 3 from the clamped state
    h = Xc[:, 1]
    assert np.allclose(Ac[:, 3], np.sqrt(2) * h / np.sqrt(h ** 4 + np.maximum(1e-8, h) ** 4), rtol=1e-14)


def test_refresh_off_passes_the_derivative_slots_through():
    mesh, X, A = _line_case()
    orc = oracle("sme0_1d")
    Xc, Ac = X.copy(), A.copy()
    orc.enforce(mesh, orc.default_params, Xc, Ac, refresh_derivative_aux=False)
    assert np.array_equal(Ac[:, :3], A[:, :3])
