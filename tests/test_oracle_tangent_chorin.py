"""Pins the tangent-linear Chorin oracle (liboracle_vam_tl.so), the reference gradient of the
Chorin adjoint, on the CPU.

  * its primal loss is bitwise the double oracle's (same operations, dual parts aside);
  * its directional derivative equals central differences of the double oracle's loss, to
    1e-6 relative, at order 1 and at order 2 (exact limiter), with the pressure solved to
    machine precision so that the differences see the exact solve;
  * linearity in the direction.
The pressure's tangent is that of the exact solution of A(X) P = -R0(X), so these checks pin
the linearised pressure equation (dA P + dR0 through the residual's own transcription) and the
predictor / corrector tangents."""
import numpy as np
import pytest

from conftest import TangentOracle, cases, oracle, zdm


def chorin_case(n=5, steps=4, seed=3):
    """C3-like VAM state on a sloped bed: neighbouring bed values differ by far more than the
    finite-difference step, so max(b_L, b_R) of the hydrostatic reconstruction stays on one
    side of its kink (a flat bed puts every face on the kink, where no derivative exists)."""
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    X = cases.c3_state(mesh)
    xy = mesh.cell_centroid[:, :2]
    eta = X[:, 0] + X[:, 1]
    X[:, 0] = 0.15 * xy[:, 0] + 0.05 * xy[:, 1] + 0.02 * np.sin(5 * xy[:, 0] + 3 * xy[:, 1])
    X[:, 1] = eta - X[:, 0]
    X[:, 3] = 0.05 * np.sin(3 * xy[:, 0])
    X[:, 4] = 0.1 * np.cos(2 * xy[:, 1])
    X[:, 5] = 0.02 * xy[:, 0] * xy[:, 1]
    X[:, 6] = 0.01 * np.sin(xy[:, 1])
    X[:, 7] = -0.01 * xy[:, 0]
    A = np.zeros((mesh.ncells, 1))
    face_bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    rng = np.random.default_rng(seed)
    gauges = rng.choice(mesh.ncells, 6, replace=False).astype(np.int32)
    obs = 1.0 + 0.05 * rng.standard_normal((steps // 2) * len(gauges))
    dts = np.full(steps, 2e-3)
    return mesh, face_bc, X, A, gauges, obs, dts


def direction(mesh, seed):
    rng = np.random.default_rng(seed)
    dX = np.zeros((mesh.ncells, 8))
    dX[:, 0] = 0.01 * rng.standard_normal(mesh.ncells)
    dX[:, 1] = 0.01 * rng.standard_normal(mesh.ncells)
    dX[:, 2] = 0.01 * rng.standard_normal(mesh.ncells)
    dX[:, 3] = 0.005 * rng.standard_normal(mesh.ncells)
    dX[:, 6] = 0.005 * rng.standard_normal(mesh.ncells)
    return dX


@pytest.mark.parametrize("order", [1, 2])
def test_tangent_chorin_loss_vs_central_differences(order):
    orc, tl = oracle("vam"), TangentOracle("vam")
    mesh, fbc, X, A, gauges, obs, dts = chorin_case()
    p = orc.default_params
    L0 = orc.chorin_loss(mesh, fbc, p, X, A, dts, gauges, 2, obs, order=order)
    for seed in (1, 2):
        dX = direction(mesh, seed)
        L, dL = tl.chorin_loss(mesh, fbc, p, X, dX, A, None, dts, gauges, 2, obs, order=order)
        assert L == L0
        eps = 1e-6
        Lp = orc.chorin_loss(mesh, fbc, p, X + eps * dX, A, dts, gauges, 2, obs, order=order)
        Lm = orc.chorin_loss(mesh, fbc, p, X - eps * dX, A, dts, gauges, 2, obs, order=order)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(dL) > 1e-6
        assert abs(fd - dL) <= 1e-6 * abs(dL), (fd, dL)


def test_tangent_chorin_is_linear_in_the_direction():
    orc, tl = oracle("vam"), TangentOracle("vam")
    mesh, fbc, X, A, gauges, obs, dts = chorin_case(steps=2)
    p = orc.default_params
    d1, d2 = direction(mesh, 1), direction(mesh, 2)
    _, a = tl.chorin_loss(mesh, fbc, p, X, d1, A, None, dts, gauges, 2, obs)
    _, b = tl.chorin_loss(mesh, fbc, p, X, d2, A, None, dts, gauges, 2, obs)
    _, c = tl.chorin_loss(mesh, fbc, p, X, 2.0 * d1 - 3.0 * d2, A, None, dts, gauges, 2, obs)
    assert abs(c - (2.0 * a - 3.0 * b)) <= 1e-9 * (abs(a) + abs(b))
