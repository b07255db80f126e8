"""The device fully implicit step (time_integration "implicit": BDF2 on transport + source,
Newton-Krylov with a matrix-free Jacobian and block-Jacobi preconditioning) against the
oracle's, which solves the same BDF equation to round-off by Newton with a dense Jacobian
(tests/test_oracle_bdf.py pins it): with tight Newton / Krylov tolerances the device states
agree to 1e-9 after several steps (BDF1 start, then variable-step BDF2), at order 1 and at
order 2 (unlimited reconstruction, a smooth residual), and the water column is conserved to
round-off.  With the reference's default tolerances the solves converge in a few Newton
iterations and the water column drifts only by the Newton tolerance."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu

TIGHT = dict(snes_atol=1e-30, snes_rtol=1e-13, snes_max_it=30, ksp_rtol=1e-13, ksp_max_it=3000)


def slip_case(n=6, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed, renumber=True)
    X = cases.c1_state(mesh, 6, seed=seed)
    xy = mesh.cell_centroid[:, :2]
    X[:, 2] = 0.3 * np.sin(2 * np.pi * xy[:, 1])
    X[:, 3] = 0.1 * np.sin(2 * np.pi * xy[:, 0])
    X[:, 4] = 0.2 * np.cos(2 * np.pi * xy[:, 0])
    return mesh, X, np.zeros((mesh.ncells, 1))


@pytest.mark.parametrize("order", [1, 2])
def test_implicit_steps_match_the_oracle_solution(order):
    mesh, X, A = slip_case()
    print("case implicit order", order, "cells", mesh.ncells, flush=True)
    s = zdm.Solver("sme1", mesh, order=order, cfl=0.4, t_end=1e9, time_integration="implicit",
                   use_limiters=False)
    p = s.params.copy()
    p[1], p[2] = 0.5, 0.01
    s.close()
    s = zdm.Solver("sme1", mesh, params=p, order=order, cfl=0.4, t_end=1e9, time_integration="implicit",
                   use_limiters=False)
    s.implicit_setup(**TIGHT)
    s.upload(X, A)
    s.initialize()
    s.step(4)
    Xg, _ = s.download()
    newton, ksp, last, reason = s.implicit_stats()
    assert newton >= 4 and ksp > 0 and reason in (2, 3, -5)
    orc = oracle("sme1")
    Xo, Ao = X.copy(), A.copy()
    n, t, dt = orc.run(mesh, s.face_bc, p, Xo, Ao, order=order, use_limiters=0, cfl=0.4, max_steps=4,
                       time_integration="implicit")
    tg, dtg, steps = s.time()
    assert n == 4 and steps == 4
    assert abs(tg - t) <= 1e-9 * t
    assert np.abs(Xg - Xo).max() <= 1e-9 * np.abs(Xo).max(), np.abs(Xg - Xo).max()
    vol = mesh.cell_volume
    assert abs(np.dot(Xg[:, 1], vol) - np.dot(X[:, 1], vol)) <= 1e-12 * np.dot(X[:, 1], vol)
    s.close()


@pytest.mark.parametrize("order,lim", [(1, True), (2, False)])
def test_implicit_default_tolerances_converge_and_conserve_mass(order, lim):
    """With the reference's SNES / KSP defaults every step converges in a few Newton
    iterations (order 1, and order 2 with a smooth, unlimited reconstruction)."""
    mesh, X, A = slip_case(n=12)
    s = zdm.Solver("sme1", mesh, order=order, cfl=0.45, t_end=1e9, time_integration="implicit",
                   use_limiters=lim)
    s.upload(X, A)
    s.initialize()
    for _ in range(5):
        s.step(1)
        newton, ksp, last, reason = s.implicit_stats()
        assert reason in (2, 3) and 1 <= last <= 4
    Xg, _ = s.download()
    assert ksp >= newton
    # the water column is conserved to the Newton tolerance (sum vol G_h = sum vol Xdot_h once
    # the fluxes cancel): atol 1e-4 on |G| bounds the drift per step by dt |G|
    vol = mesh.cell_volume
    tg, _, _ = s.time()
    assert abs(np.dot(Xg[:, 1], vol) - np.dot(X[:, 1], vol)) <= 5 * 1e-4 * tg * np.sqrt(vol.sum())
    assert np.all(np.isfinite(Xg))
    s.close()


def test_implicit_with_the_limiter_keeps_the_linear_solves_cheap():
    """Order 2 with the limiter: the matrix-free Jacobian holds the face limiter values of the
    Newton base point fixed, so each Newton system takes a handful of GMRES iterations
    (differencing across the limiter's switches -- local extrema, alpha = 1 for |dq| < 1e-12,
    one alpha for all components -- took thousands).  The limited residual itself jumps at
    those switches, so Newton stops at max_it short of rtol 1e-3 here, as any Newton on it
    would; the state stays finite and the water column within the unconverged residual."""
    mesh, X, A = slip_case(n=12)
    s = zdm.Solver("sme1", mesh, order=2, cfl=0.45, t_end=1e9, time_integration="implicit", use_limiters=True)
    s.upload(X, A)
    s.initialize()
    s.step(3)
    Xg, _ = s.download()
    newton, ksp, last, reason = s.implicit_stats()
    assert newton >= 3 and ksp <= 20 * newton, (newton, ksp)
    assert np.all(np.isfinite(Xg))
    s.close()


def test_implicit_step_is_not_captured():
    mesh, X, A = slip_case()
    s = zdm.Solver("sme1", mesh, order=1, cfl=0.4, t_end=1e9, time_integration="implicit")
    s.upload(X, A)
    s.initialize()
    with pytest.raises(zdm.ZdmError):
        s.step_graph(1)
    s.close()
