"""The device IMEX step (time_integration "imex", ARK2: explicit transport, implicit source)
against the oracle's, which tests/test_oracle_imex.py pins (tableau, stability function, time
order): bitwise over several steps at order 1 and 2, with the Manning friction (Newton stage
solves) and with MOOD; the water column to round-off."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu


def case(model, n=10, seed=5):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed, renumber=True)
    info = zdm.ModelInfo(model)
    X = cases.c1_state(mesh, info.n_q, seed=seed)
    xy = mesh.cell_centroid[:, :2]
    X[:, 2] = 0.3 * np.sin(2 * np.pi * xy[:, 1])
    X[:, 4] = 0.2 * np.cos(2 * np.pi * xy[:, 0])
    A = np.zeros((mesh.ncells, info.n_aux))
    if model == "sme1m":
        A[:, 1] = 0.03 + 0.02 * xy[:, 0]
    return mesh, X, A


@pytest.mark.parametrize("model,order", [("sme1", 1), ("sme1", 2), ("sme1m", 1), ("sme1m", 2), ("sme2", 2)])
def test_imex_steps_bitwise_vs_oracle(model, order):
    mesh, X, A = case(model)
    print("case imex", model, "order", order, "cells", mesh.ncells, flush=True)
    s = zdm.Solver(model, mesh, order=order, ssp_stages=2, cfl=0.4, t_end=1e9, time_integration="imex")
    s.upload(X, A)
    s.initialize()
    s.step(6)
    Xg, Ag = s.download()
    orc = oracle(model)
    Xo, Ao = X.copy(), A.copy()
    n, t, dt = orc.run(mesh, s.face_bc, s.params, Xo, Ao, order=order, ssp_stages=2, cfl=0.4, max_steps=6,
                       time_integration="imex")
    tg, dtg, steps = s.time()
    assert n == 6 and steps == 6
    assert tg == t and dtg == dt
    assert np.array_equal(Xg, Xo)
    vol = mesh.cell_volume
    assert abs(np.dot(Xg[:, 1], vol) - np.dot(X[:, 1], vol)) <= 1e-13 * np.dot(X[:, 1], vol)
    s.close()


def test_imex_with_mood_bitwise_vs_oracle():
    mesh, X, A = case("sme1")
    X[::7, 1] = -0.5   # holes: troubled cells for the MOOD override
    s = zdm.Solver("sme1", mesh, order=2, ssp_stages=2, cfl=0.4, t_end=1e9, time_integration="imex",
                   positivity="mood")
    s.upload(X, A)
    s.initialize()
    s.step(3)
    Xg, _ = s.download()
    orc = oracle("sme1")
    Xo, Ao = X.copy(), A.copy()
    orc.run(mesh, s.face_bc, s.params, Xo, Ao, order=2, ssp_stages=2, cfl=0.4, max_steps=3,
            time_integration="imex", positivity="mood")
    assert orc.troubled > 0 and s.troubled_cells() == orc.troubled
    assert np.array_equal(Xg, Xo)
    s.close()


def test_imex_graph_replay_matches_eager_steps():
    mesh, X, A = case("sme1m")
    out = []
    for graph in (False, True):
        s = zdm.Solver("sme1m", mesh, order=2, ssp_stages=2, cfl=0.4, t_end=1e9, time_integration="imex")
        s.upload(X, A)
        s.initialize()
        for _ in range(3):
            if graph:
                s.step_graph(2)
            else:
                s.step(2)
            s.synchronize()
        out.append(s.download()[0])
        s.close()
    assert np.array_equal(out[0], out[1])
