"""The C5 calibration problem (zoomy_dmplex_amd.cases.c5_case: dam break over a smooth random
bed with Manning friction, 32 gauges every 5 steps, first guess from other bed and Manning
fields) at test size, 50 steps on the binomial schedule with its stretch tape: the device
adjoint against the tangent-linear oracle (forward-mode dual numbers through the same
discrete map) on smooth and white-noise directions in (bed, Manning n), in both limiter
modes.  This separates the adjoint's correctness from the smoothness of the loss: central
differences of this problem stop agreeing with any linearisation after a few steps, because
the limiter's ratios of small slopes change at the finite-difference scale."""
import numpy as np
import pytest

from conftest import TangentOracle, cases, zdm

pytestmark = pytest.mark.gpu

MODEL = "sme1m"


def directions(mesh, k, seed=21):
    rng = np.random.default_rng(seed)
    srng = cases.SplitMix64(seed)
    xy = mesh.cell_centroid[:, :2]
    out = []
    for i in range(k):
        if i % 2 == 0:
            db, dn = 0.01 * cases.smooth_field(xy, srng), 0.005 * cases.smooth_field(xy, srng)
        else:
            db, dn = 0.01 * rng.standard_normal(mesh.ncells), 0.005 * rng.standard_normal(mesh.ncells)
        out.append((db, dn))
    return out


# The tangent-linear oracle differentiates the source's Newton iterations as they run (the
# reference's stop at |R| < 1e-8, SourceStep.hpp:125-209), the adjoint the converged solve
# (implicit function theorem): the derivative of the last iterate differs from the converged
# one by O(|R| of the iterate before it) per cell and step, which 50 steps of Manning friction
# accumulate to about 1e-5.  Without the source both differentiate the same operations.
TOL = {False: 1e-9, True: 1e-4}


@pytest.mark.parametrize("source", [False, True])
@pytest.mark.parametrize("exact", [True, False])
def test_c5_problem_adjoint_vs_tangent_linear(exact, source):
    mesh = zdm.Mesh.tri_grid(24, 24, jitter=0.15, seed=5)
    Xt, At, Xg, Ag, gauges = cases.c5_case(mesh)
    T, oe = 50, 5
    s = zdm.Solver(MODEL, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=source, use_limiters=True)
    s.adjoint_setup(T, gauges=gauges, obs_every=oe)
    s.adjoint_forward(Xt, At)
    obs = s.adjoint_residuals().copy()
    s.adjoint_setup(T, gauges=gauges, obs_every=oe, obs=obs, checkpoint_every=-2 - 8, exact_limiter=exact)
    assert s.adjoint_stretch_tape()[0] == 7
    print("case c5 problem exact", exact, "source", source, "cells", mesh.ncells, flush=True)
    s.adjoint_forward(Xg, Ag)
    dts = s.adjoint_dts()
    gX, gA = s.adjoint_gradient()
    assert np.all(np.isfinite(gX)) and np.all(np.isfinite(gA))
    tl = TangentOracle(MODEL)
    for db, dn in directions(mesh, 4):
        dX, dA = np.zeros_like(Xg), np.zeros_like(Ag)
        dX[:, 0], dA[:, 1] = db, dn
        _, dL = tl.loss(mesh, s.face_bc, s.params, Xg, dX, Ag, dA, dts, gauges, oe, obs, order=2,
                        freeze_limiter=not exact, source=source)
        ad = float(np.dot(gX[:, 0], db) + np.dot(gA[:, 1], dn))
        print("adjoint vs tangent-linear rel err", abs(ad - dL) / abs(dL), flush=True)
        assert abs(ad - dL) <= TOL[source] * abs(dL), (ad, dL)
    s.close()
