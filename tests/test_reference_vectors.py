"""The restated generated headers against vectors of the reference's own expressions.

tests/golden/reference_expr_vectors.json holds inputs and outputs of the reference's generated
per-point kernels, evaluated from their expression text (ChorinOps.H:100-138, Numerics.H:32-107,
Model.H:94-274 and 356-375; see reference_expr_vectors.py, which compiles and imports nothing of
the reference).  The product's restatements of those headers (models/vam/ChorinOps.H,
models/sme0_1d/Model.H and Numerics.H) are evaluated at the same inputs through the oracle
library, which is compiled from the same model headers as the device code.

Tolerance: the restatements regroup the generated sums (ChorinOps.H factors out dt / rho,
pow(x, n) becomes products, pow(x, 1/2) becomes sqrt), so each output may differ from the text's
evaluation by rounding: |ours - ref| <= 1e-14 * S, where S is the largest magnitude of that
output and of the terms it is built from over the whole vector set (an output that cancels to
nearly zero is compared on the scale of its terms, not of its tiny value).  Identical expression
structure (boundary ghost, clamp, initial condition, zero entries) is compared bit for bit.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import ORACLE_DIR, ROOT

VEC = os.path.join(ROOT, "tests", "golden", "reference_expr_vectors.json")
_dp = ctypes.POINTER(ctypes.c_double)


def dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def vectors():
    with open(VEC) as f:
        return json.load(f)


def lib(model):
    L = ctypes.CDLL(os.path.join(ORACLE_DIR, "liboracle_%s.so" % model))
    return L


def close(ours, ref, scale, tol=1e-14):
    ours, ref = np.asarray(ours, float), np.asarray(ref, float)
    return np.all(np.abs(ours - ref) <= tol * scale), np.max(np.abs(ours - ref) / scale)


def test_vam_pressure_blocks_residual_corrector(vectors):
    L = lib("vam")
    L.orc_vam_point.argtypes = [_dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
    rows = vectors["vam"]
    assert len(rows) >= 40
    ours = {"blocks": [], "residual": [], "corrector": []}
    for r in rows:
        b, res, cor = np.zeros(20), np.zeros(2), np.zeros(6)
        Q, A, p = (np.array(r[k], float) for k in ("Q", "Qaux", "p"))
        L.orc_vam_point(dp(Q), dp(A), dp(p), r["dt"], b.ctypes.data_as(_dp), res.ctypes.data_as(_dp),
                        cor.ctypes.data_as(_dp))
        ours["blocks"].append(b)
        ours["residual"].append(res)
        ours["corrector"].append(cor)
    for key in ours:
        mine = np.array(ours[key])
        ref = np.array([r[key] for r in rows])
        # scale per output: the largest |value| of the output and of the state / aux magnitudes
        # times dt / rho that enter it (the generated sums have terms of that size)
        terms = np.array([max(np.max(np.abs(r["Q"])), np.max(np.abs(r["Qaux"]))) ** 3 *
                          max(1.0, r["dt"] / r["p"][1]) for r in rows])
        scale = np.maximum(np.abs(ref).max(axis=0)[None, :], terms[:, None])
        ok, worst = close(mine, ref, scale)
        assert ok, (key, worst)
        # the zero entries of the generated blocks are exact zeros here too
        assert np.array_equal(mine == 0, ref == 0) or key != "blocks", key


def test_sme0_1d_numerical_flux_fluctuations_eigenvalue(vectors):
    L = lib("sme0_1d")
    L.orc_flux.argtypes = [_dp] * 8
    L.orc_eigen.argtypes = [_dp] * 5
    rows = vectors["sme0_1d_numerics"]
    assert len(rows) >= 60
    F, D, E = [], [], []
    for r in rows:
        f, d, e = np.zeros(3), np.zeros(6), np.zeros(1)
        args = [np.array(r[k], float) for k in ("Q_minus", "Q_plus", "Qaux_minus", "Qaux_plus", "p", "n")]
        L.orc_flux(*[dp(a) for a in args], f.ctypes.data_as(_dp), d.ctypes.data_as(_dp))
        L.orc_eigen(dp(args[0]), dp(args[2]), dp(args[4]), dp(args[5]), e.ctypes.data_as(_dp))
        F.append(f)
        D.append(d)
        E.append(e)
    for mine, key in ((F, "flux"), (D, "fluctuations"), (E, "lmax_minus")):
        mine = np.array(mine)
        ref = np.array([r[key] for r in rows])
        # per sample: the flux-sized scale of that face (momentum flux magnitudes, including the
        # near-dry faces where hu^2 / h is large)
        scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-300)
        ok, worst = close(mine, ref, scale)
        assert ok, (key, worst)
        assert np.array_equal(mine == 0, ref == 0), key


def test_sme0_1d_model_functions(vectors):
    L = lib("sme0_1d")
    L.orc_model_point.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int] + \
                                 [_dp] * 8
    rows = vectors["sme0_1d_model"]
    keys = ["source", "source_jacobian_wrt_variables", "source_jacobian_wrt_aux_variables",
            "update_variables", "update_aux_variables", "eigenvalues", "boundary_conditions",
            "initial_condition"]
    sizes = [3, 9, 12, 3, 4, 3, 3, 3]
    ours = {k: [] for k in keys}
    for r in rows:
        outs = [np.zeros(n) for n in sizes]
        Q, A, p, n, X = (np.array(r[k], float) for k in ("Q", "Qaux", "p", "n", "X"))
        L.orc_model_point(dp(Q), dp(A), dp(p), dp(n), dp(X), r["time"], r["dt"], 0,
                          *[o.ctypes.data_as(_dp) for o in outs])
        for k, o in zip(keys, outs):
            ours[k].append(o)
    exact = {"source_jacobian_wrt_aux_variables", "update_variables", "boundary_conditions",
             "initial_condition"}
    for k in keys:
        mine = np.array(ours[k])
        ref = np.array([r[k] for r in rows])
        if k in exact:
            assert np.array_equal(mine, ref), k
            continue
        scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-300)
        ok, worst = close(mine, ref, scale)
        assert ok, (k, worst)
        assert np.array_equal(mine == 0, ref == 0), k


def test_vectors_cover_dry_and_wet_states(vectors):
    hs = [r["Q_minus"][1] for r in vectors["sme0_1d_numerics"]] + \
         [r["Q_plus"][1] for r in vectors["sme0_1d_numerics"]]
    assert min(hs) < 0 and any(h == 0 for h in hs) and any(0 < h < 1e-8 for h in hs)
    assert sum(h > 0.02 for h in hs) > 60
    # both normal orientations of the 1-D face and a non-unit one
    ns = {r["n"][0] for r in vectors["sme0_1d_numerics"]}
    assert 1.0 in ns and -1.0 in ns and len(ns) > 2
