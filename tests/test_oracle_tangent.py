"""Pins the tangent-linear oracle (the reference gradient of the discrete adjoint) and the
Manning friction model it runs on (CPU).

  * Manning source: the implicit Euler update satisfies (q1 - q0)/dt = S(q1) with the closed
    form S_hu = -g n^2 |u| u h^(-1/3) (|u| regularised by 1e-8);
  * the tangent-linear run's loss equals the plain oracle's loss bitwise (same arithmetic on
    the values), and its directional derivative matches central finite differences of the
    plain oracle with frozen time steps.
The case has a sloped bed, so neighbouring bed values differ by far more than the
finite-difference step: max(b_L, b_R) of the hydrostatic reconstruction and the limiter's
local-extremum switches stay on one side of their kinks (where they do not, finite
differences straddle a kink and no derivative exists to compare)."""
import numpy as np
import pytest

from conftest import TangentOracle, cases, oracle, zdm

MODEL = "sme1m"


def calib_case(n=6, steps=8, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    rng = np.random.default_rng(seed)
    X = cases.c1_state(mesh, 6, seed=seed)
    xy = mesh.cell_centroid[:, :2]
    eta = X[:, 0] + X[:, 1]
    X[:, 0] = 0.15 * xy[:, 0] + 0.05 * xy[:, 1] + 0.02 * np.sin(5 * xy[:, 0] + 3 * xy[:, 1])
    X[:, 1] = eta - X[:, 0]
    X[:, 2] = 0.05 * rng.standard_normal(mesh.ncells)
    X[:, 4] = 0.05 * rng.standard_normal(mesh.ncells)
    A = np.zeros((mesh.ncells, 2))
    A[:, 1] = rng.uniform(0.01, 0.05, mesh.ncells)
    bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    gauges = rng.choice(mesh.ncells, 5, replace=False).astype(np.int32)
    dts = np.full(steps, 2e-3)
    obs = 0.9 + 0.05 * rng.standard_normal((steps // 2, len(gauges)))
    return mesh, X, A, bc, gauges, dts, obs


def test_manning_source_closed_form():
    orc = oracle(MODEL)
    mesh, X, A, *_ = calib_case()
    X0 = X.copy()
    dt = 0.05
    orc.source(orc.default_params, dt, X, A)
    h = X[:, 1]
    u, v = X[:, 2] / h, X[:, 4] / h
    n = A[:, 1]
    S = -9.81 * n ** 2 * np.sqrt(u * u + v * v + 1e-16) * h ** (-1.0 / 3.0)
    # Newton stops at |R| < 1e-8 (SourceStep.hpp:161): the update is exact to 1e-8 / dt
    assert np.allclose((X[:, 2] - X0[:, 2]) / dt, S * u, rtol=1e-6, atol=1e-8 / dt)
    assert np.allclose((X[:, 4] - X0[:, 4]) / dt, S * v, rtol=1e-6, atol=1e-8 / dt)
    assert np.sum(X[:, 2] != X0[:, 2]) > mesh.ncells // 2
    assert np.array_equal(X[:, [0, 1, 3, 5]], X0[:, [0, 1, 3, 5]])


@pytest.mark.parametrize("order", [1, 2])
def test_tangent_linear_matches_central_differences(order):
    orc, tl = oracle(MODEL), TangentOracle(MODEL)
    mesh, X, A, bc, gauges, dts, obs = calib_case()
    p = orc.default_params
    L0 = orc.loss(mesh, bc, p, X, A, dts, gauges, 2, obs, order=order)
    rng = np.random.default_rng(7)
    for _ in range(4):
        dX = np.zeros_like(X)
        dX[:, 0] = 0.01 * rng.standard_normal(mesh.ncells)     # bathymetry
        dX[:, 1] = 0.01 * rng.standard_normal(mesh.ncells)     # depth
        dA = np.zeros_like(A)
        dA[:, 1] = 0.005 * rng.standard_normal(mesh.ncells)    # Manning n
        Lt, dL = tl.loss(mesh, bc, p, X, dX, A, dA, dts, gauges, 2, obs, order=order)
        assert Lt == L0
        eps = 1e-6
        Lp = orc.loss(mesh, bc, p, X + eps * dX, A + eps * dA, dts, gauges, 2, obs, order=order)
        Lm = orc.loss(mesh, bc, p, X - eps * dX, A - eps * dA, dts, gauges, 2, obs, order=order)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - dL) <= 1e-6 * abs(dL) + 1e-12, (fd, dL)
