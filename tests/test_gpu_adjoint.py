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


def make_layout(order, steps, ck, stretch, exact):
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=6, steps=steps)
    s = zdm.Solver(MODEL, mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True)
    s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs, checkpoint_every=ck, exact_limiter=exact,
                    stretch_tape=stretch)
    return s, X, A


# (checkpoint_every, stretch tape: None automatic / 0 none / L / (L, buffers), expected layout,
# expected stretch tape (steps, level, buffers)); 8 steps
LAYOUTS_8 = [(-2, None, (1, 2), (0, 0, 0)), (-1, None, (1, 1), (0, 0, 0)), (0, None, None, None),
             (1, None, (1, 0), (0, 0, 0)), (3, None, (3, 0), (3, 2, 2)), (3, 0, (3, 0), (0, 0, 0)),
             (3, (3, 1), (3, 0), (3, 2, 1)), (8, None, (8, 0), (8, 2, 2)), (8, 0, (8, 0), (0, 0, 0)),
             (-3, None, (-3, 0), (8, 2, 2)), (-3, 0, (-3, 0), (0, 0, 0)), (-3, 3, (-3, 0), (3, 2, 2)),
             (-4, None, (-4, 0), (4, 2, 2)), (-4, 0, (-4, 0), (0, 0, 0)), (-4, (1, 1), (-4, 0), (1, 2, 1)),
             (-4, (1, 2), (-4, 0), (1, 2, 2)), (-5, None, (-5, 0), (3, 2, 2)), (-5, 0, (-5, 0), (0, 0, 0)),
             (-5, (2, 1), (-5, 0), (2, 2, 1)), (-7, None, (-7, 0), (2, 2, 2)), (-7, 0, (-7, 0), (0, 0, 0))]


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("order", [1, 2])
def test_trajectory_layout_does_not_change_the_gradient(order, exact):
    """Tape with face states (-2), tape of stage states (-1), every state (1), checkpoints
    (3, 8), the binomial schedule on 1, 2, 3 and 5 slots (-3, -4, -5, -7), each with its
    automatic stretch tape, without one and with one of another length, and the automatic
    choice (0): bitwise the same gradient, since what a tape keeps is what the re-runs
    recompute.  In the frozen-limiter mode (exact False) the full tape with face states
    carries the stored limiter values and the other layouts recompute them."""
    grads = []
    for ck, stretch, want, want_st in LAYOUTS_8:
        print("case order", order, "exact", exact, "checkpoint_every", ck, "stretch", stretch, flush=True)
        s, X, A = make_layout(order, 8, ck, stretch, exact)
        if want is not None:
            assert s.adjoint_layout() == want
            assert s.adjoint_stretch_tape() == want_st
        s.adjoint_forward(X, A)
        grads.append(s.adjoint_gradient())
        s.close()
    assert np.all(np.isfinite(grads[0][0])) and np.abs(grads[0][0]).max() > 0
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and np.array_equal(g[1], grads[0][1])


@pytest.mark.parametrize("exact", [True, False])
def test_binomial_schedule_on_a_longer_trajectory(exact):
    """21 steps on 3 and 4 slots (split by the dynamic programme, nested several levels deep),
    with and without stretch tapes (2 and 3 steps leave the slots too few to cover the
    trajectory, so stretches are reversed in chunks from the scratch state), with one tape
    buffer and with two (re-runs on the second stream), and checkpoints every 4 with a stretch
    tape: bitwise the gradient of the tape."""
    grads = []
    for ck, stretch in ((-2, None), (-5, 0), (-5, None), (-5, 2), (-5, (2, 1)), (-6, 0), (-6, None), (-6, 3),
                        (-6, (3, 2)), (4, None), (4, (4, 1))):
        print("case exact", exact, "checkpoint_every", ck, "stretch", stretch, flush=True)
        s, X, A = make_layout(2, 21, ck, stretch, exact)
        s.adjoint_forward(X, A)
        grads.append(s.adjoint_gradient())
        s.close()
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and np.array_equal(g[1], grads[0][1])


def test_stretch_tape_is_refused_where_it_does_not_apply():
    s, X, A = make_layout(2, 8, 1, None, True)
    with pytest.raises(zdm.ZdmError, match="stretch"):
        s.adjoint_stretch_tape(3)
    s.close()
    s, X, A = make_layout(2, 8, -2, None, True)
    with pytest.raises(zdm.ZdmError, match="full tape"):
        s.adjoint_stretch_tape(3)
    s.close()


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


def test_device_entry_points_and_autograd_streams():
    """zdm_adjoint_forward_device / zdm_adjoint_gradient_device on torch device buffers give
    bitwise the host entry points' loss and gradient; the autograd Function gives them too when
    torch works on a side stream (ordered through the solver stream) and from CPU tensors."""
    torch = pytest.importorskip("torch")
    import importlib
    ag = importlib.import_module("zoomy_dmplex_amd.autograd")
    s, mesh, X, A, bc, gauges, obs = make(2)
    L_h = s.adjoint_forward(X, A)
    gX, gA = s.adjoint_gradient()
    Xd = torch.tensor(X, dtype=torch.float64, device="cuda")
    Ad = torch.tensor(A, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert s.adjoint_forward_device(Xd.data_ptr(), Ad.data_ptr()) == L_h
    gXd = torch.full_like(Xd, float("nan"))
    gAd = torch.full_like(Ad, float("nan"))
    torch.cuda.synchronize()
    s.adjoint_gradient_device(gXd.data_ptr(), gAd.data_ptr())
    s.synchronize()
    assert np.array_equal(gXd.cpu().numpy(), gX) and np.array_equal(gAd.cpu().numpy(), gA)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        X0 = torch.tensor(X, dtype=torch.float64, device="cuda").requires_grad_(True)
        A0 = torch.tensor(A, dtype=torch.float64, device="cuda").requires_grad_(True)
        L = ag.calibration_loss(s, X0, A0)
        (2.0 * L).backward()
        gx2, ga2 = (X0.grad * 0.5).cpu().numpy(), (A0.grad * 0.5).cpu().numpy()
    side.synchronize()
    assert float(L) == L_h
    assert np.array_equal(gx2, gX) and np.array_equal(ga2, gA)
    X1 = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    A1 = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    ag.calibration_loss(s, X1, A1).backward()
    assert np.array_equal(X1.grad.numpy(), gX) and np.array_equal(A1.grad.numpy(), gA)
    s.close()


def test_forward_that_reaches_t_end_early_is_an_error():
    """Steps past t_end are no-ops of the forward with nothing recorded: refused."""
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=6, steps=8)
    s = zdm.Solver(MODEL, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e-9, use_limiters=True)
    s.adjoint_setup(8, gauges=gauges, obs_every=2, obs=obs)
    with pytest.raises(zdm.ZdmError, match="t_end"):
        s.adjoint_forward(X, A)
    s.close()


@pytest.mark.parametrize("source", [False, True])
def test_frozen_limiter_gradient_vs_differences_of_the_frozen_map(source):
    """The frozen-limiter gradient is the derivative of the forward map whose face limiter
    values are held at the unperturbed run's (zdm_adjoint_forward_frozen_alpha): central
    differences of that map agree to 1e-4 over 30 steps, against about 1e-2 and more for
    differences of the plain map (whose limiter values move with the perturbation).  The
    residual is the frozen map's own kinks: where Barth-Jespersen clips a face value to a
    neighbour's value, that neighbour's own face value there is often the same number (a local
    extremum), so bstar = max(bL, bR) ties at the base point; the plain map moves both sides
    together, the frozen one separates them, and central differences there average two
    one-sided slopes (the gradient takes one of them, as the forward's branch does).  The
    recording run itself is the plain forward: same loss and dt sequence."""
    steps = 30
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=10, steps=steps)
    s = zdm.Solver(MODEL, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True, source=source)
    s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs, checkpoint_every=-7, exact_limiter=False)
    L0 = s.adjoint_forward(X, A)
    dts0 = s.adjoint_dts().copy()
    gX, gA = s.adjoint_gradient()
    Lr = s.adjoint_forward_frozen_alpha(X, A, record=True)
    assert Lr == L0 and np.array_equal(s.adjoint_dts(), dts0)
    assert s.adjoint_forward_frozen_alpha(X, A, dts=dts0) == L0   # the replay at the base point
    eps = 1e-6
    frozen, plain = [], []
    for dX, dA in directions(mesh, X, A, 4, seed=17):
        ad = np.sum(gX * dX) + np.sum(gA * dA)
        fd = (s.adjoint_forward_frozen_alpha(X + eps * dX, A + eps * dA, dts=dts0) -
              s.adjoint_forward_frozen_alpha(X - eps * dX, A - eps * dA, dts=dts0)) / (2 * eps)
        frozen.append(abs(ad - fd) / abs(fd))
    for dX, dA in directions(mesh, X, A, 4, seed=17):
        ad = np.sum(gX * dX) + np.sum(gA * dA)
        fd = (s.adjoint_forward(X + eps * dX, A + eps * dA, dts=dts0) -
              s.adjoint_forward(X - eps * dX, A - eps * dA, dts=dts0)) / (2 * eps)
        plain.append(abs(ad - fd) / abs(fd))
    print("frozen gradient vs differences: frozen map", frozen, "plain map", plain, flush=True)
    assert max(frozen) <= 1e-4, frozen
    assert np.median(plain) >= 10 * np.median(frozen), (plain, frozen)
    s.close()
