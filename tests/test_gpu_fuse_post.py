"""The final SSP stage fused with the post-step (k_flux_post) against the separate kernels
(k_flux<MODE_FINAL> then k_post): the same operations on the same values, so states, aux,
time and dt are bitwise equal, with the split source (constraints, Newton, constraints) and
without it, one and several cells per thread, plain and graph-replayed steps, and an
ensemble whose members take their own dt."""
import numpy as np
import pytest

from conftest import cases, zdm

pytestmark = pytest.mark.gpu


def _run(model, mesh, X, A, steps, fuse, launch=None, graph=False, **kw):
    s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=1, **kw)
    s.fuse_post(fuse)
    if launch:
        s.set_launch(*launch)
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


@pytest.mark.parametrize("launch,graph", [(None, False), ((64, 2), False), ((128, 1), True)])
def test_fused_post_with_source_matches_separate_kernels(launch, graph):
    mesh = zdm.Mesh.tri_grid(40, 40, jitter=0.1, seed=21)
    X = cases.c2_state(mesh, 8, seed=2)
    rng = np.random.default_rng(21)
    X[:, 2:] = 0.05 * rng.standard_normal((mesh.ncells, 6))
    A = np.zeros((mesh.ncells, 1))
    print("case sme2 + source, launch", launch, "graph", graph, flush=True)
    a = _run("sme2", mesh, X, A, 30, True, launch, graph, source=True)
    b = _run("sme2", mesh, X, A, 30, False, launch, graph, source=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    assert np.max(np.abs(a[0][:, 3])) > 0


def test_fused_post_without_source_matches_separate_kernels():
    mesh = zdm.Mesh.tri_grid(30, 30, jitter=0.15, seed=22)
    X = cases.c1_state(mesh, 6, seed=22)
    A = np.zeros((mesh.ncells, zdm.ModelInfo("sme1").n_aux))
    print("case sme1 without source", flush=True)
    a = _run("sme1", mesh, X, A, 40, True)
    b = _run("sme1", mesh, X, A, 40, False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_fused_post_in_an_ensemble_matches_separate_kernels():
    info = zdm.ModelInfo("sme1")
    ms, Xs = [], []
    for i, n in enumerate((5, 11, 17, 8)):
        m = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=60 + i)
        X = cases.c1_state(m, info.n_q, seed=60 + i)
        X[:, 1] *= 0.6 + 0.1 * i
        ms.append(m)
        Xs.append(X)
    out = []
    for fuse in (True, False):
        print("case ensemble, fuse", fuse, flush=True)
        e = zdm.Ensemble("sme1", ms, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
        e.fuse_post(fuse)
        e.upload(e.join(Xs))
        e.initialize()
        e.step(20)
        Xe, Ae = e.download()
        out.append((Xe, Ae, [e.member_time(i) for i in range(len(ms))]))
        e.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("ck", [-2, -6])
def test_fused_post_in_the_adjoint_matches_separate_kernels(ck):
    """The adjoint's recording step fuses its final stage with the recording post-stage
    (k_flux_post<REC>: X1, the aux before the source and the source's solution kept as
    k_post_record keeps them): loss, dt sequence and gradient bitwise those of the separate
    kernels, on the full tape and on the binomial schedule with its stretch tape."""
    from test_oracle_tangent import calib_case
    steps = 12
    mesh, X, A, bc, gauges, dts, obs = calib_case(n=10, steps=steps)
    out = []
    for fuse in (True, False):
        print("case adjoint, checkpoint_every", ck, "fuse", fuse, flush=True)
        s = zdm.Solver("sme1m", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True,
                       source=True)
        s.fuse_post(fuse)
        s.adjoint_setup(steps, gauges=gauges, obs_every=2, obs=obs, checkpoint_every=ck)
        L = s.adjoint_forward(X, A)
        out.append((L, s.adjoint_dts().copy(), s.adjoint_gradient()))
        s.close()
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2][0], out[1][2][0]) and np.array_equal(out[0][2][1], out[1][2][1])
    assert np.abs(out[0][2][0]).max() > 0
