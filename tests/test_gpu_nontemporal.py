"""The non-temporal form of the stage and final flux launches (zdm_flux_nontemporal) against the
ordinary one: only the cache policy of some loads and stores differs, so states, aux, time and
dt are bitwise equal, fused with the post-step and not, in plain and graph-replayed steps, with
several cells per thread (where the ordinary form is kept), in an ensemble and in the adjoint's
recording step.  The default picks the form by the size of one flux pass."""
import numpy as np
import pytest

from conftest import cases, zdm

pytestmark = pytest.mark.gpu


def test_default_picks_the_form_by_size():
    small = zdm.Mesh.tri_grid(40, 40)
    s = zdm.Solver("sme2", small, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True)
    assert s.flux_nontemporal() is False
    assert s.flux_nontemporal(1) is True
    assert s.flux_nontemporal(0) is False
    s.close()
    # 583 200 cells: 583 200 x (48 x 8 + 48) + 40 x 875 880 B = 287 MB per flux pass > 256 MiB
    large = zdm.Mesh.tri_grid(540, 540)
    s = zdm.Solver("sme2", large, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True)
    assert s.flux_nontemporal() is True
    s.close()


def _run(mesh, X, A, steps, nt, fuse=True, launch=None, graph=False):
    s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=1, source=True)
    s.fuse_post(fuse)
    if launch:
        s.set_launch(*launch)
    assert s.flux_nontemporal(nt) == bool(nt)
    s.upload(X, A)
    s.initialize()
    if graph:
        s.step_graph(steps)
    else:
        s.step(steps)
    Xg, Ag = s.download()
    t = s.time()
    s.close()
    return Xg, Ag, t


@pytest.mark.parametrize("fuse,launch,graph", [(True, None, False), (False, None, False), (True, None, True),
                                               (True, (64, 2), False)])
def test_nontemporal_form_matches_ordinary_form(fuse, launch, graph):
    mesh = zdm.Mesh.tri_grid(40, 40, jitter=0.1, seed=21)
    X = cases.c2_state(mesh, 8, seed=2)
    rng = np.random.default_rng(21)
    X[:, 2:] = 0.05 * rng.standard_normal((mesh.ncells, 6))
    A = np.zeros((mesh.ncells, 1))
    print("case sme2 + source, fuse", fuse, "launch", launch, "graph", graph, flush=True)
    a = _run(mesh, X, A, 30, 1, fuse, launch, graph)
    b = _run(mesh, X, A, 30, 0, fuse, launch, graph)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    assert np.max(np.abs(a[0][:, 3])) > 0


def test_nontemporal_form_in_an_ensemble():
    info = zdm.ModelInfo("sme1")
    ms, Xs = [], []
    for i, n in enumerate((5, 11, 17, 8)):
        m = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=60 + i)
        X = cases.c1_state(m, info.n_q, seed=60 + i)
        X[:, 1] *= 0.6 + 0.1 * i
        ms.append(m)
        Xs.append(X)
    out = []
    for nt in (1, 0):
        print("case ensemble, nontemporal", nt, flush=True)
        e = zdm.Ensemble("sme1", ms, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
        e.flux_nontemporal(nt)
        e.upload(e.join(Xs))
        e.initialize()
        e.step(20)
        Xe, Ae = e.download()
        out.append((Xe, Ae, [e.member_time(i) for i in range(len(ms))]))
        e.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_nontemporal_form_in_the_adjoint_recording_step():
    from test_oracle_tangent import calib_case
    steps = 12
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=10, steps=steps)
    out = []
    for nt in (1, 0):
        print("case adjoint, nontemporal", nt, flush=True)
        s = zdm.Solver("sme1m", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True,
                       source=True)
        s.flux_nontemporal(nt)
        s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs, checkpoint_every=-6)
        L = s.adjoint_forward(X, A)
        out.append((L, s.adjoint_dts().copy(), s.adjoint_gradient()))
        s.close()
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2][0], out[1][2][0]) and np.array_equal(out[0][2][1], out[1][2][1])
    assert np.abs(out[0][2][0]).max() > 0


@pytest.mark.parametrize("ti", ["imex", "implicit"])
def test_nontemporal_form_in_imex_and_implicit_steps(ti):
    """The right-hand-side and IMEX flux launches (k_flux MODE_RHS / MODE_IMEX3) in both forms."""
    mesh = zdm.Mesh.tri_grid(16, 16, jitter=0.15, seed=31)
    X = cases.c2_state(mesh, 8, seed=31)
    rng = np.random.default_rng(31)
    X[:, 2:] = 0.05 * rng.standard_normal((mesh.ncells, 6))
    out = []
    for nt in (1, 0):
        print("case sme2", ti, "nontemporal", nt, flush=True)
        kw = dict(use_limiters=0) if ti == "implicit" else {}
        s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.4, t_end=1e9, source=True, time_integration=ti,
                       **kw)
        assert s.flux_nontemporal(nt) == bool(nt)
        s.upload(X, np.zeros((mesh.ncells, 1)))
        s.initialize()
        s.step(4)
        Xg, Ag = s.download()
        out.append((Xg, Ag, s.time()))
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]) and out[0][2] == out[1][2]
