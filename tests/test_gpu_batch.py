"""Ensemble (batching form) and graph-replayed steps against single runs and the oracle.

An ensemble member runs the same kernels on the same arithmetic as a solver built on its own
mesh, and its dt is the minimum over its own cells, so member states are bitwise equal to the
single runs.  A graph replay launches the same kernels with the same arguments as zdm_step."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu

MODEL = "sme1"


def members(sizes=(3, 7, 12, 20, 9), seed0=40):
    ms, Xs = [], []
    info = zdm.ModelInfo(MODEL)
    for i, n in enumerate(sizes):
        m = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed0 + i)
        X = cases.c1_state(m, info.n_q, seed=seed0 + i)
        X[:, 1] = X[:, 1] * (0.6 + 0.1 * i)      # different depths -> different dt
        ms.append(m)
        Xs.append(X)
    return info, ms, Xs


def single_run(mesh, X, steps, **kw):
    s = zdm.Solver(MODEL, mesh, **kw)
    s.upload(X)
    s.initialize()
    s.step(steps)
    Xg, _ = s.download()
    t = s.time()
    s.close()
    return Xg, t


KW = dict(order=2, ssp_stages=2, cfl=0.45, t_end=1e9)


def test_ensemble_members_equal_single_runs():
    info, ms, Xs = members()
    print("case ensemble", [m.ncells for m in ms], flush=True)
    e = zdm.Ensemble(MODEL, ms, **KW)
    e.upload(e.join(Xs))
    e.initialize()
    e.step(25)
    Xe, _ = e.download()
    orc = oracle(MODEL)
    for i, (m, X, Xm) in enumerate(zip(ms, Xs, e.split(Xe))):
        Xs1, (t1, dt1, n1) = single_run(m, X, 25, **KW)
        assert np.array_equal(Xm, Xs1), i
        te, dte, ne = e.member_time(i)
        assert (te, dte, ne) == (t1, dt1, n1)
        Xc, Ac = X.copy(), np.zeros((m.ncells, info.n_aux))
        orc.run(m, zdm.boundary_face_bc(m, info.boundary_tags, 2), e.params, Xc, Ac,
                order=2, ssp_stages=2, cfl=0.45, max_steps=25)
        assert np.max(np.abs(Xm - Xc)) <= 1e-12 * np.max(np.abs(Xc))
    # the members really took different time steps
    assert len({e.member_time(i)[1] for i in range(len(ms))}) == len(ms)
    e.close()


def test_ensemble_t_end_per_member():
    info, ms, Xs = members(sizes=(4, 9, 6))
    kw = dict(KW, t_end=0.01)
    e = zdm.Ensemble(MODEL, ms, **kw)
    e.upload(e.join(Xs))
    e.initialize()
    n = e.run(max_steps=500, check_every=7)
    Xe, _ = e.download()
    steps = []
    for i, (m, X, Xm) in enumerate(zip(ms, Xs, e.split(Xe))):
        s = zdm.Solver(MODEL, m, **kw)
        s.upload(X)
        s.initialize()
        s.run(max_steps=500, check_every=7)
        Xs1, _ = s.download()
        t1 = s.time()
        s.close()
        assert np.array_equal(Xm, Xs1)
        assert e.member_time(i) == t1
        assert t1[0] == 0.01
        steps.append(t1[2])
    # zdm_run reports member 0's steps; the members needed different step counts
    assert n == steps[0] and len(set(steps)) > 1
    e.close()


@pytest.mark.parametrize("ensemble", [False, True])
def test_graph_replay_equals_eager(ensemble):
    info, ms, Xs = members(sizes=(6, 11))
    if ensemble:
        make = lambda: zdm.Ensemble(MODEL, ms, **KW)
        X0 = np.concatenate(Xs)
    else:
        make = lambda: zdm.Solver(MODEL, ms[1], **KW)
        X0 = Xs[1]
    a = make()
    a.upload(X0)
    a.initialize()
    a.step(12)
    Xa, _ = a.download()
    b = make()
    b.upload(X0)
    b.initialize()
    done = []
    for k in (3, 4, 5):            # replays with eager work and synchronisation in between
        b.step_graph(k)
        Xb, _ = b.download()
        done.append(Xb)
        b.cfl_dt() if not ensemble else b.member_time(0)
    assert np.array_equal(done[-1], Xa)
    assert b.time() == a.time()
    a.close()
    b.close()
