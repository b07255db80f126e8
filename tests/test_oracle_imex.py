"""Pins the oracle's IMEX step (time_integration "imex": explicit transport + implicit source in
one ARK2 step, the additive Runge-Kutta method of Giraldo, Kelly and Constantinescu, SIAM J.
Sci. Comput. 35(5), 2013) on the CPU.

  * the published tableau (copied here from the paper) satisfies the order-1 and order-2
    conditions of both parts and the row-sum (c) conditions;
  * on the scalar test equation y' = le y + li y the oracle's step is the tableau's stability
    function R(ze, zi) = 1 + (ze b + zi bt)^T (I - ze A - zi At)^-1 1, to 1e-14 (so every stage
    coefficient, the stage derivative (Y - Z)/(dt g) and the weights are those of the paper);
  * it is L-stable in the stiff limit (R -> 0 as zi -> -inf at ze = 0);
  * on a mesh (SME level 1 with a strong slip friction, k = nu/lambda = 50, whose source is
    linear in the momenta so that the stage Newton solves are exact) the step conserves the
    water column to round-off and converges in time at second order: halving the CFL number
    quarters the error at a fixed end time.  (With a nonlinear source the stage derivative
    (Y - Z)/(dt g) carries the Newton tolerance 1e-8 / (dt g), which puts a floor under the
    error; that is the method as PETSc runs it.)
Against the reference itself the step is parity unpinned: PETSc (TSARKIMEX, SNES) is not
available, its "2e" may register another member of this family, and its SNES stops on global
residual norms (atol 1e-4, rtol 1e-3, MUSCLSolver.hpp:108-110), while the stage systems here are
solved per cell to the SourceStep Newton tolerance."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

R2 = np.sqrt(2.0)
G, D = 1.0 - 1.0 / R2, 1.0 / (2.0 * R2)
A = np.array([[0, 0, 0], [2 * G, 0, 0], [1 - (3 + 2 * R2) / 6, (3 + 2 * R2) / 6, 0]])
AT = np.array([[0, 0, 0], [G, G, 0], [D, D, G]])
B = np.array([D, D, G])
C = np.array([0, 2 * G, 1.0])


def stability(ze, zi):
    Y = np.linalg.solve(np.eye(3) - ze * A - zi * AT, np.ones(3))
    return 1.0 + (ze * B + zi * B) @ Y


def test_published_tableau_is_second_order():
    for M in (A, AT):
        assert np.allclose(M.sum(axis=1), C, atol=1e-15)
        assert abs(B.sum() - 1) < 1e-15
        assert abs(B @ C - 0.5) < 1e-15
    # coupling conditions of an ARK pair of order 2 hold since both parts share b and c
    assert np.allclose(np.diag(AT)[1:], G) and AT[0, 0] == 0


@pytest.mark.parametrize("ze,zi", [(0.3, 0.0), (0.0, -0.7), (-0.4, -2.5), (0.2, -30.0), (-1.1, -0.05)])
def test_scalar_step_is_the_stability_function(ze, zi):
    orc = oracle("sme1m")
    dt = 0.01
    for n in (1, 5):
        y = orc.ark2_linear(ze / dt, zi / dt, dt, n, 1.0)
        r = stability(ze, zi) ** n
        assert abs(y - r) <= 1e-14 * max(1.0, abs(r)), (n, y, r)


def test_l_stable_in_the_stiff_limit():
    orc = oracle("sme1m")
    for zi in (-1e4, -1e6, -1e8):
        assert abs(orc.ark2_linear(0.0, zi, 1.0, 1, 1.0)) < 10.0 / abs(zi)
    # A-stable along the negative axis
    for zi in np.linspace(-50, -0.01, 40):
        assert abs(orc.ark2_linear(0.0, zi, 1.0, 1, 1.0)) <= 1.0


def slip_case(n=12, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    X = cases.c1_state(mesh, 6, seed=seed)
    xy = mesh.cell_centroid[:, :2]
    X[:, 2] = 0.3 * np.sin(2 * np.pi * xy[:, 1])
    X[:, 3] = 0.1 * np.sin(2 * np.pi * xy[:, 0])
    X[:, 4] = 0.2 * np.cos(2 * np.pi * xy[:, 0])
    return mesh, X, np.zeros((mesh.ncells, 1))


def test_imex_conserves_mass_and_is_second_order_in_time():
    orc = oracle("sme1")
    mesh, X0, A0 = slip_case()
    bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    p = orc.default_params.copy()
    p[1], p[2] = 0.5, 0.01   # nu, lambda: k_slip = 50
    vol = mesh.cell_volume
    t_end = 0.02
    sols = {}
    for cfl in (0.4, 0.2, 0.1, 0.05, 0.0125):
        X, A = X0.copy(), A0.copy()
        steps, t, _ = orc.run(mesh, bc, p, X, A, order=1, cfl=cfl, t_end=t_end, max_steps=100000,
                              time_integration="imex")
        assert abs(t - t_end) < 1e-15
        assert abs(np.dot(X[:, 1], vol) - np.dot(X0[:, 1], vol)) <= 1e-13 * np.dot(X0[:, 1], vol)
        sols[cfl] = X
    ref = sols[0.0125]
    err = [np.abs(sols[c] - ref).max() for c in (0.4, 0.2, 0.1, 0.05)]
    rates = [np.log2(err[i] / err[i + 1]) for i in range(3)]
    assert all(r > 1.8 for r in rates), (err, rates)


def test_imex_without_a_source_is_the_explicit_ark():
    """With a zero source the IMEX step reduces to the explicit three-stage part; on the scalar
    equation that is R(ze, 0) = 1 + ze + ze^2/2 + a32 a21 b3 ze^3."""
    orc = oracle("sme1m")
    for ze in (0.1, -0.5, 0.9):
        y = orc.ark2_linear(ze, 0.0, 1.0, 1, 1.0)
        assert abs(y - (1 + ze + ze ** 2 / 2 + A[2, 1] * A[1, 0] * B[2] * ze ** 3)) < 1e-14
