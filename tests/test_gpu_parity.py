"""GPU path vs the CPU oracle on the same seeded inputs (bitwise expected: both use only
correctly-rounded +,-,*,/,sqrt with no contraction, in the same per-cell face order)."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu

MODELS = ["sme0_1d", "sme0", "sme1", "sme2"]


def make_case(model, n=10, jitter=0.15, seed=3, moments=True):
    info = zdm.ModelInfo(model)
    rng = np.random.default_rng(seed)
    if info.dim == 1:
        mesh = zdm.Mesh.line(48, 0.0, 1.0)
        x = mesh.cell_centroid[:, 0]
        X = np.zeros((mesh.ncells, info.n_q))
        X[:, 0] = 0.2 * np.exp(-(x - 0.5) ** 2 / 0.01)
        X[:, 1] = np.where(x < 0.4, 1.0, 0.5) + 1e-3 * rng.standard_normal(mesh.ncells)
        X[:, 2] = 0.1 * rng.standard_normal(mesh.ncells)
    else:
        mesh = zdm.Mesh.tri_grid(n, n, jitter=jitter, seed=seed)
        X = cases.c1_state(mesh, info.n_q, seed=seed)
        if moments:
            X[:, 2:] = 0.05 * rng.standard_normal((mesh.ncells, info.n_q - 2))
    return info, mesh, X


def rel_err(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / scale


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("order,gradient", [(2, 0), (2, 1), (1, 0)])
def test_transport_rhs_matches_oracle(model, order, gradient):
    info, mesh, X = make_case(model)
    print("case", model, "order", order, "gradient", gradient, "cells", mesh.ncells, flush=True)
    orc = oracle(model)
    s = zdm.Solver(model, mesh, order=order, gradient=gradient, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, info.n_aux))
    F_gpu = s.transport_rhs(X, A, t=0.0)
    F_cpu = orc.form_rhs(mesh, s.face_bc, s.params, X.copy(), A.copy(), order=order,
                         gradient=gradient)
    assert np.all(np.isfinite(F_cpu))
    assert rel_err(F_gpu, F_cpu) <= 1e-14, rel_err(F_gpu, F_cpu)
    s.close()


@pytest.mark.parametrize("model", MODELS)
def test_time_steps_match_oracle(model):
    info, mesh, X = make_case(model)
    print("case", model, "steps", flush=True)
    orc = oracle(model)
    s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, info.n_aux))
    s.upload(X, A)
    s.initialize()
    s.step(20)
    Xg, Ag = s.download()
    tg, dtg, ng = s.time()
    Xc, Ac = X.copy(), A.copy()
    n, tc, dtc = orc.run(mesh, s.face_bc, s.params, Xc, Ac, order=2, ssp_stages=2, cfl=0.45,
                         max_steps=20)
    assert n == ng == 20
    assert rel_err(Xg, Xc) <= 1e-12, rel_err(Xg, Xc)
    assert abs(tg - tc) <= 1e-12 * abs(tc)
    assert abs(dtg - dtc) <= 1e-12 * abs(dtc)
    s.close()


@pytest.mark.parametrize("model", ["sme0", "sme2"])
def test_ssp_stages_and_t_end_clamp(model):
    info, mesh, X = make_case(model)
    orc = oracle(model)
    s = zdm.Solver(model, mesh, order=2, ssp_stages=3, cfl=0.3, t_end=0.004)
    A = np.zeros((mesh.ncells, info.n_aux))
    s.upload(X, A)
    s.initialize()
    s.step(200)   # runs past t_end: the extra steps must be no-ops
    Xg, _ = s.download()
    tg, _, ng = s.time()
    Xc, Ac = X.copy(), A.copy()
    n, tc, _ = orc.run(mesh, s.face_bc, s.params, Xc, Ac, order=2, ssp_stages=3, cfl=0.3,
                       t_end=0.004, max_steps=200)
    assert ng == n and tg == tc == 0.004
    assert rel_err(Xg, Xc) <= 1e-12


@pytest.mark.parametrize("model", MODELS)
def test_source_and_cfl_match_oracle(model):
    info, mesh, X = make_case(model)
    orc = oracle(model)
    s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, info.n_aux))
    s.upload(X, A)
    s.enforce_constraints()
    Xe, Ae = s.download()
    Xc, Ac = X.copy(), A.copy()
    orc.enforce(mesh, s.params, Xc, Ac)   # incl. RefreshDerivativeAux for the 1-D model
    assert np.array_equal(Xe, Xc) and np.array_equal(Ae, Ac)
    dt_g = s.cfl_dt()
    dt_c = orc.compute_dt(mesh, s.params, Xc, Ac, 0.45)
    assert dt_g == dt_c
    s.source_step(0.05)
    Xs, _ = s.download()
    orc.source(s.params, 0.05, Xc, Ac)
    assert rel_err(Xs, Xc) <= 1e-14


@pytest.mark.parametrize("model", ["sme1", "sme2", "sme1m"])
@pytest.mark.parametrize("dt", [0.05, 10.0, 1e3])
def test_source_newton_with_row_exchanges_matches_oracle(model, dt):
    """Large source steps make |dt dS/dh| exceed the unit pivot of the depth row, so GEPP
    exchanges rows in some cells (the device's exchange path) and not in others."""
    info, mesh, X = make_case(model)
    rng = np.random.default_rng(17)
    X[:, 2:] = 0.3 * rng.standard_normal((mesh.ncells, info.n_q - 2))
    s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, info.n_aux))
    if model == "sme1m":
        A[:, 1] = 0.03
    s.upload(X, A)
    s.enforce_constraints()
    Xe, Ae = s.download()
    print("case", model, "source dt", dt, flush=True)
    s.source_step(dt)
    Xs, _ = s.download()
    s.close()
    Xc, Ac = Xe.copy(), Ae.copy()
    oracle(model).source(s.params, dt, Xc, Ac)
    assert np.all(np.isfinite(Xc))
    assert np.array_equal(Xs, Xc), rel_err(Xs, Xc)


def test_c1_mass_and_parity_200_steps():
    """C1 (SME-L0 dam break over a bump, 60x60 grid): 200 SSP-RK2 steps vs the oracle."""
    mesh = zdm.Mesh.tri_grid(60, 60)
    X = cases.c1_state(mesh, 4, seed=1)
    s = zdm.Solver("sme0", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, 1))
    s.upload(X, A)
    s.initialize()
    s.step(200)
    Xg, _ = s.download()
    orc = oracle("sme0")
    Xc, Ac = X.copy(), A.copy()
    orc.run(mesh, s.face_bc, s.params, Xc, Ac, order=2, ssp_stages=2, cfl=0.45, max_steps=200)
    assert rel_err(Xg, Xc) <= 1e-12, rel_err(Xg, Xc)
    m0 = np.sum(mesh.cell_volume * X[:, 1])
    m1 = np.sum(mesh.cell_volume * Xg[:, 1])
    assert abs(m1 - m0) <= 1e-13 * m0


@pytest.mark.parametrize("model", ["sme0", "sme1", "sme2"])
@pytest.mark.parametrize("order,wb", [(1, 0), (2, 1)])
def test_lake_at_rest(model, order, wb):
    """Order 1 (PCM) is exactly well-balanced in the reference scheme; order 2 needs the
    centred hydrostatic term (well_balanced extension)."""
    info = zdm.ModelInfo(model)
    mesh = zdm.Mesh.tri_grid(24, 24, jitter=0.15, seed=11)
    X = cases.lake_at_rest(mesh, info.n_q)
    s = zdm.Solver(model, mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, well_balanced=wb)
    s.upload(X, None)
    s.initialize()
    s.step(50)
    Xg, _ = s.download()
    eta = Xg[:, 0] + Xg[:, 1]
    assert np.max(np.abs(eta - 1.0)) <= 1e-13
    assert np.max(np.abs(Xg[:, 2:])) <= 1e-13
    Xc, Ac = X.copy(), np.zeros((mesh.ncells, info.n_aux))
    oracle(model).run(mesh, s.face_bc, s.params, Xc, Ac, order=order, ssp_stages=2, cfl=0.45,
                      max_steps=50, well_balanced=wb)
    assert rel_err(Xg, Xc) <= 1e-12


@pytest.mark.parametrize("model", ["sme1", "sme2"])
def test_well_balanced_rhs_matches_oracle(model):
    info, mesh, X = make_case(model)
    s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, well_balanced=1)
    A = np.zeros((mesh.ncells, info.n_aux))
    F_gpu = s.transport_rhs(X, A)
    F_cpu = oracle(model).form_rhs(mesh, s.face_bc, s.params, X.copy(), A.copy(), well_balanced=1)
    assert rel_err(F_gpu, F_cpu) <= 1e-14
