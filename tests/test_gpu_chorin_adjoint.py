"""Discrete adjoint of the Chorin pressure split on the device (VAM), with the pressure solves
run to machine precision so that the forward is the exact split:

  * <gradient, v> against the tangent-linear oracle's directional derivative (the oracle's own
    linearised pressure equation, pinned by central differences on the CPU), order 1 and
    order 2 with the exact limiter, to 1e-8 relative;
  * against central differences of the device's own frozen-dt forward, to 1e-6 relative;
  * every state layout (all states, checkpoints, binomial schedules) gives bitwise the same
    gradient."""
import numpy as np
import pytest

from conftest import TangentOracle, oracle, zdm
from test_oracle_tangent_chorin import chorin_case, direction

pytestmark = pytest.mark.gpu

KRYLOV = dict(restart=30, max_it=3000, rtol=1e-13, atol=1e-30, op_mode=1)


def make(order, steps=4, ck=0, exact=True):
    mesh, fbc, X, A, gauges, obs, dts = chorin_case(steps=steps)
    s = zdm.Solver("vam", mesh, order=order, cfl=0.45, t_end=1e9, use_limiters=True, face_bc=fbc)
    s.chorin_setup(**KRYLOV)
    s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs.reshape(steps // 2, len(gauges)),
                    checkpoint_every=ck, exact_limiter=exact, chorin=True)
    return s, mesh, fbc, X, A, gauges, obs, dts


@pytest.mark.parametrize("order", [1, 2])
def test_chorin_adjoint_vs_tangent_linear(order):
    s, mesh, fbc, X, A, gauges, obs, dts = make(order)
    print("case chorin adjoint vs TL, order", order, "cells", mesh.ncells, flush=True)
    loss = s.adjoint_forward(X, A, dts=dts)
    gX, _ = s.adjoint_gradient()
    orc, tl = oracle("vam"), TangentOracle("vam")
    p = s.params
    L0 = orc.chorin_loss(mesh, fbc, p, X, A, dts, gauges, 2, obs, order=order)
    assert abs(loss - L0) <= 1e-11 * abs(L0)
    assert s.adjoint_krylov_its() > 0
    for seed in (1, 2, 3):
        dX = direction(mesh, seed)
        _, dL = tl.chorin_loss(mesh, fbc, p, X, dX, A, None, dts, gauges, 2, obs, order=order)
        ad = float(np.sum(gX * dX))
        assert abs(ad - dL) <= 1e-8 * abs(dL), (seed, ad, dL)
    s.close()


def test_chorin_adjoint_vs_device_central_differences():
    s, mesh, fbc, X, A, gauges, obs, dts = make(1)
    print("case chorin adjoint vs device FD, cells", mesh.ncells, flush=True)
    s.adjoint_forward(X, A, dts=dts)
    gX, _ = s.adjoint_gradient()
    for seed in (4, 5):
        dX = direction(mesh, seed)
        eps = 1e-6
        Lp = s.adjoint_forward(X + eps * dX, A, dts=dts)
        Lm = s.adjoint_forward(X - eps * dX, A, dts=dts)
        fd = (Lp - Lm) / (2 * eps)
        ad = float(np.sum(gX * dX))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (seed, ad, fd)
    s.close()


def test_chorin_adjoint_layouts_agree_bitwise():
    grads = []
    for ck, want in ((1, (1, 0)), (3, (3, 0)), (-3, (-3, 0)), (-4, (-4, 0)), (0, None)):
        s, mesh, fbc, X, A, gauges, obs, dts = make(2, steps=6, ck=ck, exact=False)
        print("case chorin adjoint layout", ck, flush=True)
        if want is not None:
            assert s.adjoint_layout() == want
        s.adjoint_forward(X, A)
        grads.append(s.adjoint_gradient()[0])
        s.close()
    assert np.all(np.isfinite(grads[0])) and np.abs(grads[0]).max() > 0
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def test_chorin_adjoint_rejects_tapes():
    mesh, fbc, X, A, gauges, obs, dts = chorin_case(steps=4)
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9, face_bc=fbc)
    with pytest.raises(zdm.ZdmError):
        s.adjoint_setup(4, gauges=gauges, obs_every=2, obs=obs.reshape(2, len(gauges)), checkpoint_every=-2,
                        chorin=True)
    s.close()
