"""Discrete adjoint on the device against the tangent-linear oracle and finite differences.

The gradient is the transpose of the derivative of the discrete step with frozen branches
and time steps, so <gradient, v> must equal the tangent-linear oracle's directional
derivative along v (same discrete map, forward mode); central differences of the device's
own frozen-dt forward check it independently."""
import numpy as np
import pytest

from conftest import TangentOracle, cases, oracle, zdm
from test_oracle_tangent import calib_case

pytestmark = pytest.mark.gpu

MODEL = "sme1m"


def make(order, steps=8, n=6, ck=0, exact=True):
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=n, steps=steps)
    s = zdm.Solver(MODEL, mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True)
    s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs, checkpoint_every=ck,
                    exact_limiter=exact)
    return s, mesh, X, A, bc, gauges, obs


def directions(mesh, X, A, k, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        dX = np.zeros_like(X)
        dX[:, 0] = 0.01 * rng.standard_normal(mesh.ncells)
        dX[:, 1] = 0.01 * rng.standard_normal(mesh.ncells)
        dX[:, 2] = 0.01 * rng.standard_normal(mesh.ncells)
        dA = np.zeros_like(A)
        dA[:, 1] = 0.005 * rng.standard_normal(mesh.ncells)
        out.append((dX, dA))
    return out


@pytest.mark.parametrize("order,exact", [(1, True), (2, True), (2, False)])
def test_adjoint_loss_and_gradient_vs_tangent_linear(order, exact):
    s, mesh, X, A, bc, gauges, obs = make(order, exact=exact)
    print("case adjoint order", order, "cells", mesh.ncells, flush=True)
    L = s.adjoint_forward(X, A)
    dts = s.adjoint_dts()
    assert np.all(dts > 0)
    orc = oracle(MODEL)
    Lo = orc.loss(mesh, bc, s.params, X, A, dts, gauges, 2, obs, order=order)
    assert abs(L - Lo) <= 1e-13 * abs(Lo)
    gX, gA = s.adjoint_gradient()
    assert np.all(np.isfinite(gX)) and np.all(np.isfinite(gA))
    tl = TangentOracle(MODEL)
    for dX, dA in directions(mesh, X, A, 4):
        _, dL = tl.loss(mesh, bc, s.params, X, dX, A, dA, dts, gauges, 2, obs, order=order,
                        freeze_limiter=not exact)
        ad = np.sum(gX * dX) + np.sum(gA * dA)
        # the tangent run differentiates the Newton iterations (finite-difference aux Jacobian,
        # stop at |R| < 1e-8); the adjoint transposes the converged implicit step
        print("adjoint vs tangent-linear rel err", abs(ad - dL) / abs(dL))
        assert abs(ad - dL) <= 1e-6 * abs(dL) + 1e-14, (ad, dL)
    s.close()


def test_adjoint_vs_central_differences_on_device():
    s, mesh, X, A, bc, gauges, obs = make(2)
    s.adjoint_forward(X, A)
    dts = s.adjoint_dts()
    gX, gA = s.adjoint_gradient()
    eps = 1e-6
    for dX, dA in directions(mesh, X, A, 3, seed=9):
        Lp = s.adjoint_forward(X + eps * dX, A + eps * dA, dts=dts)
        Lm = s.adjoint_forward(X - eps * dX, A - eps * dA, dts=dts)
        fd = (Lp - Lm) / (2 * eps)
        ad = np.sum(gX * dX) + np.sum(gA * dA)
        print("adjoint vs central differences rel err", abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    s.close()


@pytest.mark.parametrize("order", [1, 2])
def test_trajectory_layout_does_not_change_the_gradient(order):
    """Tape with face states (-2), tape of stage states (-1), every state (1), checkpoints
    (3, 8), the binomial schedule on 1, 2, 3 and 5 slots (-3, -4, -5, -7) and the automatic
    choice (0): bitwise the same gradient, since what a tape keeps is what the re-runs
    recompute."""
    grads = []
    for ck, want in ((-2, (1, 2)), (-1, (1, 1)), (0, None), (1, (1, 0)), (3, (3, 0)), (8, (8, 0)),
                     (-3, (-3, 0)), (-4, (-4, 0)), (-5, (-5, 0)), (-7, (-7, 0))):
        s, mesh, X, A, bc, gauges, obs = make(order, ck=ck)
        if want is not None:
            assert s.adjoint_layout() == want
        s.adjoint_forward(X, A)
        grads.append(s.adjoint_gradient())
        s.close()
    assert np.all(np.isfinite(grads[0][0])) and np.abs(grads[0][0]).max() > 0
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and np.array_equal(g[1], grads[0][1])


def test_binomial_schedule_on_a_longer_trajectory():
    """21 steps on 3 and 4 slots (split by the dynamic programme, nested several levels deep):
    bitwise the gradient of the tape."""
    grads = []
    for ck in (-2, -5, -6):
        s, mesh, X, A, bc, gauges, obs = make(2, steps=21, ck=ck)
        s.adjoint_forward(X, A)
        grads.append(s.adjoint_gradient())
        s.close()
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and np.array_equal(g[1], grads[0][1])


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    import importlib
    ag = importlib.import_module("zoomy_dmplex_amd.autograd")
    s, mesh, X, A, bc, gauges, obs = make(2)
    X0 = torch.tensor(X, dtype=torch.float64, device="cuda", requires_grad=True)
    A0 = torch.tensor(A, dtype=torch.float64, device="cuda", requires_grad=True)
    L = ag.calibration_loss(s, X0, A0)
    (3.0 * L).backward()
    gX, gA = s.adjoint_gradient()
    assert np.array_equal(X0.grad.cpu().numpy(), 3.0 * gX)
    assert np.array_equal(A0.grad.cpu().numpy(), 3.0 * gA)
    # directional finite difference through the same Function with the recorded time steps
    dts = s.adjoint_dts()
    rng = np.random.default_rng(2)
    db = torch.tensor(0.01 * rng.standard_normal(mesh.ncells), dtype=torch.float64, device="cuda")
    dn = torch.tensor(0.005 * rng.standard_normal(mesh.ncells), dtype=torch.float64, device="cuda")
    eps = 1e-6
    with torch.no_grad():
        Xp, Xm, Ap, Am = X0.clone(), X0.clone(), A0.clone(), A0.clone()
        Xp[:, 0] += eps * db
        Xm[:, 0] -= eps * db
        Ap[:, 1] += eps * dn
        Am[:, 1] -= eps * dn
        fd = (ag.calibration_loss(s, Xp, Ap, dts) - ag.calibration_loss(s, Xm, Am, dts)) / (2 * eps)
    ad = float(torch.sum(X0.grad[:, 0] * db) + torch.sum(A0.grad[:, 1] * dn)) / 3.0
    assert abs(ad - float(fd)) <= 1e-6 * abs(float(fd))
    s.close()


def test_forward_that_reaches_t_end_early_is_an_error():
    """Steps past t_end are no-ops of the forward with nothing recorded: refused."""
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=6, steps=8)
    s = zdm.Solver(MODEL, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e-9, use_limiters=True)
    s.adjoint_setup(8, gauges=gauges, obs_every=2, obs=obs)
    with pytest.raises(zdm.ZdmError, match="t_end"):
        s.adjoint_forward(X, A)
    s.close()
