"""The optional parts of the generated-model surface on the device, against the oracle.

- RefreshDerivativeAux (TransportStep.hpp:161-202) through the 1-D SME level 0 model's
  field-level update_aux_variables (Model.H:250-262): every stage's FormRHS and every
  EnforcePhysicalConstraints refresh the derivative aux; state and aux bitwise equal to the
  oracle, the derivative slots non-trivial, graph replays equal to eager steps, and with
  solver.refresh_derivative_aux off the slots pass through.
- Model::diffusive_flux (TransportStep.hpp:468-491, 533-550) of the sme0_visc model: the
  transport right-hand side and 20 steps against the oracle at orders 1 and 2 with both
  gradients; the term changes the result (nu_h = 0 differs).
"""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu


def line_case(n=64, seed=9):
    rng = np.random.default_rng(seed)
    mesh = zdm.Mesh.line(n, 0.0, 1.0)
    x = mesh.cell_centroid[:, 0]
    X = np.zeros((n, 3))
    X[:, 0] = 0.2 * np.exp(-(x - 0.5) ** 2 / 0.01)
    X[:, 1] = np.where(x < 0.4, 1.0, 0.5) + 1e-3 * rng.standard_normal(n)
    X[:, 2] = 0.1 * rng.standard_normal(n)
    return mesh, X


@pytest.mark.parametrize("order", [1, 2])
def test_refresh_derivative_aux_matches_oracle(order):
    mesh, X = line_case()
    s = zdm.Solver("sme0_1d", mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, refresh_derivative_aux=True)
    A = np.zeros((mesh.ncells, 4))
    s.upload(X, A)
    s.initialize()
    s.step(20)
    Xg, Ag = s.download()
    s.close()
    Xc, Ac = X.copy(), A.copy()
    n, _, _ = oracle("sme0_1d").run(mesh, s.face_bc, s.params, Xc, Ac, order=order, ssp_stages=2, cfl=0.45,
                                    max_steps=20, refresh_derivative_aux=True)
    assert n == 20
    assert np.array_equal(Xg, Xc)
    assert np.array_equal(Ag, Ac)
    assert np.max(np.abs(Ag[:, 1])) > 1e-2 and np.max(np.abs(Ag[:, 0])) > 1e-2   # dh/dx, d(hu)/dx


def test_refresh_derivative_aux_off_passes_slots_through():
    mesh, X = line_case()
    A = np.zeros((mesh.ncells, 4))
    A[:, :3] = np.random.default_rng(3).standard_normal((mesh.ncells, 3))
    s = zdm.Solver("sme0_1d", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, refresh_derivative_aux=False)
    s.upload(X, A)
    s.initialize()
    s.step(10)
    Xg, Ag = s.download()
    s.close()
    assert np.array_equal(Ag[:, :3], A[:, :3])
    Xc, Ac = X.copy(), A.copy()
    oracle("sme0_1d").run(mesh, s.face_bc, s.params, Xc, Ac, order=2, ssp_stages=2, cfl=0.45, max_steps=10,
                          refresh_derivative_aux=False)
    assert np.array_equal(Xg, Xc) and np.array_equal(Ag, Ac)


def test_refresh_derivative_aux_graph_replay_equals_eager():
    mesh, X = line_case()
    out = []
    for graph in (False, True):
        s = zdm.Solver("sme0_1d", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
        s.upload(X, np.zeros((mesh.ncells, 4)))
        s.initialize()
        for _ in range(3):
            s.step_graph(5) if graph else s.step(5)
            s.synchronize()
        out.append(s.download())
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_refresh_is_refused_where_the_step_does_not_carry_it():
    mesh, X = line_case()
    with pytest.raises(zdm.ZdmError, match="refresh_derivative_aux"):
        zdm.Solver("sme0_1d", mesh, order=2, positivity="mood", refresh_derivative_aux=True)
    with pytest.raises(zdm.ZdmError, match="refresh_derivative_aux"):
        zdm.Solver("sme0_1d", mesh, order=2, time_integration="imex", refresh_derivative_aux=True)


def visc_case(n=12, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    X = cases.c1_state(mesh, 4, seed=seed)
    rng = np.random.default_rng(seed)
    X[:, 2:] = 0.1 * rng.standard_normal((mesh.ncells, 2))
    return mesh, X


@pytest.mark.parametrize("order,gradient", [(2, 0), (2, 1), (1, 0)])
def test_diffusive_flux_rhs_and_steps_match_oracle(order, gradient):
    mesh, X = visc_case()
    orc = oracle("sme0_visc")
    s = zdm.Solver("sme0_visc", mesh, order=order, gradient=gradient, ssp_stages=2, cfl=0.45, t_end=1e9)
    A = np.zeros((mesh.ncells, 1))
    F_gpu = s.transport_rhs(X, A)
    F_cpu = orc.form_rhs(mesh, s.face_bc, s.params, X.copy(), A.copy(), order=order, gradient=gradient)
    assert np.array_equal(F_gpu, F_cpu)
    s.upload(X, A)
    s.initialize()
    s.step(20)
    Xg, _ = s.download()
    s.close()
    Xc, Ac = X.copy(), A.copy()
    orc.run(mesh, s.face_bc, s.params, Xc, Ac, order=order, ssp_stages=2, cfl=0.45, gradient=gradient,
            max_steps=20)
    assert np.array_equal(Xg, Xc)
    # the term acts at order 2 (it needs the cell gradients) and not at order 1
    p0 = s.params.copy()
    p0[3] = 0.0
    s0 = zdm.Solver("sme0_visc", mesh, params=p0, order=order, gradient=gradient, ssp_stages=2, cfl=0.45,
                    t_end=1e9)
    F0 = s0.transport_rhs(X, A)
    s0.close()
    if order >= 2:
        assert np.max(np.abs(F0 - F_gpu)) > 1e-8
    else:
        assert np.array_equal(F0, F_gpu)


def test_diffusive_model_refuses_the_adjoint():
    mesh, X = visc_case()
    s = zdm.Solver("sme0_visc", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    with pytest.raises(zdm.ZdmError, match="diffusive_flux"):
        s.adjoint_setup(4, gauges=np.array([0], dtype=np.int32), obs_every=2, obs=np.zeros(2))
    s.close()
