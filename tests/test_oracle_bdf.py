"""Pins the oracle's fully implicit step (time_integration "implicit": BDF2 on transport +
source, TSBDF of FullyImplicitStrategy) on the CPU.

  * the variable-step coefficients are the derivative at t_{n+1} of the Lagrange interpolant:
    exact on quadratics for BDF2 (and on linears for the BDF1 start), for unequal steps;
  * each step's nonlinear equation is solved to round-off (|G| small against |X| / dt);
  * on a mesh (SME level 1, strong slip friction k = 50) the water column is conserved and the
    error at a fixed end time is second order in the CFL number (BDF1 start included);
  * the stiff limit is damped: with a friction far beyond the step's stability limit the
    momenta decay instead of oscillating.
Against the reference the step is parity unpinned (PETSc TSBDF / SNES / KSP unavailable; the
device solves the same equation by Newton-Krylov to its SNES tolerances)."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm


@pytest.mark.parametrize("h,h1", [(0.1, 0.1), (0.1, 0.25), (0.3, 0.07)])
def test_bdf_coefficients_differentiate_the_interpolant(h, h1):
    orc = oracle("sme1")
    t1, t0, tm = 1.3, 1.3 - h, 1.3 - h - h1
    c = orc.bdf_coeffs(h, h1, True)
    for p, dp in ((lambda t: 1.0, lambda t: 0.0), (lambda t: t, lambda t: 1.0), (lambda t: t * t, lambda t: 2 * t)):
        assert abs(c[0] * p(t1) + c[1] * p(t0) + c[2] * p(tm) - dp(t1)) < 1e-12 / h
    c1 = orc.bdf_coeffs(h, h1, False)
    assert c1[2] == 0 and abs(c1[0] * t1 + c1[1] * t0 - 1.0) < 1e-12 / h


def slip_case(n=6, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    X = cases.c1_state(mesh, 6, seed=seed)
    xy = mesh.cell_centroid[:, :2]
    X[:, 2] = 0.3 * np.sin(2 * np.pi * xy[:, 1])
    X[:, 3] = 0.1 * np.sin(2 * np.pi * xy[:, 0])
    X[:, 4] = 0.2 * np.cos(2 * np.pi * xy[:, 0])
    return mesh, X, np.zeros((mesh.ncells, 1))


def test_bdf_solves_each_step_and_conserves_mass():
    orc = oracle("sme1")
    mesh, X0, A0 = slip_case()
    bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    p = orc.default_params.copy()
    p[1], p[2] = 0.5, 0.01
    vol = mesh.cell_volume
    for steps in (1, 2, 4):
        X, A = X0.copy(), A0.copy()
        _, _, dt = orc.run(mesh, bc, p, X, A, order=1, cfl=0.4, max_steps=steps, time_integration="implicit")
        r, its = orc.bdf_last_residual()
        # |G| ~ |X| / dt at the start of a step: solved to round-off of that
        assert its >= 2 and r < 1e-11 * np.abs(X).max() / dt, (steps, r, its)
        assert abs(np.dot(X[:, 1], vol) - np.dot(X0[:, 1], vol)) <= 1e-13 * np.dot(X0[:, 1], vol)


def test_bdf_is_second_order_in_time():
    orc = oracle("sme1")
    mesh, X0, A0 = slip_case()
    bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    p = orc.default_params.copy()
    p[1], p[2] = 0.5, 0.01
    t_end = 0.02
    sols = {}
    for cfl in (0.4, 0.2, 0.1, 0.025):
        X, A = X0.copy(), A0.copy()
        _, t, _ = orc.run(mesh, bc, p, X, A, order=1, cfl=cfl, t_end=t_end, max_steps=100000,
                          time_integration="implicit")
        assert abs(t - t_end) < 1e-15
        sols[cfl] = X
    ref = sols[0.025]
    err = [np.abs(sols[c] - ref).max() for c in (0.4, 0.2, 0.1)]
    rates = [np.log2(err[i] / err[i + 1]) for i in range(2)]
    assert all(r > 1.7 for r in rates), (err, rates)


def test_bdf_damps_a_stiff_friction():
    """Shallow water (SME level 0), flat bed, flat surface, a sheared current: no pressure
    forcing, a slip friction with dt k ~ 1e2 at CFL 0.4.  An L-stable step leaves a small
    fraction of the momentum; an unstable or merely A-stable one would keep it or flip its sign
    back and forth."""
    orc = oracle("sme0")
    mesh = zdm.Mesh.tri_grid(6, 6, jitter=0.15, seed=5)
    xy = mesh.cell_centroid[:, :2]
    X0 = np.zeros((mesh.ncells, 4))
    X0[:, 1] = 1.0
    X0[:, 2] = 0.3 * np.sin(2 * np.pi * xy[:, 1])
    A0 = np.zeros((mesh.ncells, orc.n_aux))
    bc = zdm.boundary_face_bc(mesh, ["default"], 2)
    p = orc.default_params.copy()
    p[1], p[2] = 50.0, 1e-3   # k_slip = 5e4
    X, A = X0.copy(), A0.copy()
    orc.run(mesh, bc, p, X, A, order=1, cfl=0.4, max_steps=3, time_integration="implicit")
    assert np.abs(X[:, 2:]).max() < 0.02 * np.abs(X0[:, 2]).max()
