"""The headline workload's own inputs on the GPU: SME level 2 + implicit Newtonian-slip source,
MUSCL with the Vacuum + Barth-Jespersen limiter, Green-Gauss gradient, SSP-RK2, CFL 0.45,
over the seeded 16-bump bed with the circular dam (cases.c2_state).

- On a 64 x 64 right-triangle grid (the 1000 x 1000 grid's shape at 8 192 cells) the device
  runs the full 500 steps with the autotuned launch shape and must match the oracle to 1e-12
  relative in every conserved variable, with the mass conserved to 1e-13.
- On the full 1000 x 1000 grid (2 000 000 cells) the device runs the 500 steps in chunks and
  checks what holds at any size: the mass to round-off, h > 0, a finite state, a strictly
  increasing time and a positive dt.
"""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu

STEPS = 500


def _mass(mesh, X):
    return float(np.sum(mesh.cell_volume * X[:, 1]))


def test_c2_inputs_500_steps_match_oracle():
    mesh = zdm.Mesh.tri_grid(64, 64)
    X = cases.c2_state(mesh, 8, seed=2)
    A = np.zeros((mesh.ncells, 1))
    s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=1,
                   source=True)
    s.upload(X, A)
    s.initialize()
    bs, cpt, _ = s.autotune(steps=3)
    print("c2 64x64: autotuned launch %d x %d, %d steps" % (bs, cpt, STEPS), flush=True)
    s.step(STEPS)
    Xg, _ = s.download()
    tg, dtg, ng = s.time()
    s.close()
    Xc, Ac = X.copy(), A.copy()
    n, tc, dtc = oracle("sme2").run(mesh, s.face_bc, s.params, Xc, Ac, order=2, ssp_stages=2,
                                    cfl=0.45, max_steps=STEPS)
    assert n == ng == STEPS
    # every conserved variable, each against its own scale
    for i in range(1, 8):
        scale = max(np.max(np.abs(Xc[:, i])), 1e-300)
        err = np.max(np.abs(Xg[:, i] - Xc[:, i])) / scale
        assert err <= 1e-12, (i, err)
    assert np.array_equal(Xg[:, 0], X[:, 0])          # the bed is not evolved
    assert abs(tg - tc) <= 1e-12 * tc and abs(dtg - dtc) <= 1e-12 * dtc
    m0, m1 = _mass(mesh, X), _mass(mesh, Xg)
    assert abs(m1 - m0) <= 1e-13 * m0, (m0, m1)
    # the run moved: the dam spread and the slip friction acted on the moments
    assert np.max(np.abs(Xg[:, 2])) > 1e-3 and np.max(np.abs(Xg[:, 4])) > 1e-5


def test_c2_full_size_500_steps_properties():
    mesh = zdm.Mesh.tri_grid(1000, 1000)
    assert mesh.ncells == 2_000_000
    X = cases.c2_state(mesh, 8, seed=2)
    s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=1,
                   source=True)
    s.upload(X, np.zeros((mesh.ncells, 1)))
    s.initialize()
    bs, cpt, _ = s.autotune(steps=3)
    print("c2 1000x1000: autotuned launch %d x %d" % (bs, cpt), flush=True)
    m0 = _mass(mesh, X)
    t_prev, done = 0.0, 0
    for chunk in (1, 49, 150, 300):
        s.step(chunk)
        t, dt, n = s.time()
        done += chunk
        assert n == done and t > t_prev and dt > 0.0, (n, t, dt)
        t_prev = t
        print("  %d steps, t = %.6g, dt = %.3g" % (n, t, dt), flush=True)
    Xg, Ag = s.download()
    s.close()
    assert np.all(np.isfinite(Xg)) and np.all(np.isfinite(Ag))
    assert np.min(Xg[:, 1]) > 0.0
    m1 = _mass(mesh, Xg)
    assert abs(m1 - m0) <= 1e-13 * m0, (m0, m1, (m1 - m0) / m0)
    assert np.array_equal(Xg[:, 0], X[:, 0])
