"""Launch shapes of the transport kernels (workgroup size x cells per thread) and the
autotuner: every shape gives bitwise the same run (single mesh, ensemble, MOOD, graph replay),
and the autotuner leaves the state where it found it."""
import numpy as np
import pytest

from conftest import cases, zdm

pytestmark = pytest.mark.gpu

SHAPES = [(256, 1), (128, 1), (64, 1), (256, 2), (64, 4), (128, 3)]


def state(mesh, n_q, seed):
    X = cases.c1_state(mesh, n_q, seed=seed)
    X[:, 2:] = 0.05 * np.random.default_rng(seed).standard_normal((mesh.ncells, n_q - 2))
    return X


@pytest.mark.parametrize("positivity", ["none", "mood"])
def test_single_mesh_shapes_bitwise_equal(positivity):
    model = "sme1"
    info = zdm.ModelInfo(model)
    mesh = zdm.Mesh.tri_grid(30, 30, jitter=0.15, seed=4)
    X = state(mesh, info.n_q, 4)
    ref = None
    for bs, cpt in SHAPES:
        print("case", positivity, bs, cpt, flush=True)
        s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, positivity=positivity)
        s.set_launch(bs, cpt)
        assert s.launch() == (bs, cpt)
        s.upload(X)
        s.initialize()
        s.step(6)
        s.step_graph(4)
        out = (s.download(), s.time())
        s.close()
        if ref is None:
            ref = out
        assert np.array_equal(out[0][0], ref[0][0]) and np.array_equal(out[0][1], ref[0][1])
        assert out[1] == ref[1]


def test_ensemble_shapes_bitwise_equal_and_autotune_restores_state():
    model = "sme1"
    info = zdm.ModelInfo(model)
    meshes = [zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=n) for n in (9, 23, 14, 31)]
    e = zdm.Ensemble(model, meshes, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    X = e.join([state(m, info.n_q, i) for i, m in enumerate(meshes)])
    ref = None
    for bs, cpt in SHAPES:
        e.set_launch(bs, cpt)
        e.upload(X)
        e.initialize()
        e.step(5)
        out = e.download()[0], [e.member_time(i) for i in range(len(meshes))]
        if ref is None:
            ref = out
        assert np.array_equal(out[0], ref[0]) and out[1] == ref[1]
    # the autotuner times every candidate and restores state, aux and time
    e.upload(X)
    e.initialize()
    e.step(2)
    before = e.download(), [e.member_time(i) for i in range(len(meshes))]
    bs, cpt, ms = e.autotune(2)
    assert (bs, cpt) in ms and len(ms) == 9 and all(v > 0 for v in ms.values())
    assert e.launch() == (bs, cpt) and ms[(bs, cpt)] == min(ms.values())
    after = e.download(), [e.member_time(i) for i in range(len(meshes))]
    assert np.array_equal(before[0][0], after[0][0]) and before[1] == after[1]
    e.step(3)
    e.close()


def test_kernel_timing_every_kth_step():
    """zdm_enable_timing(s, k): events on every k-th step of zdm_step only (the per-kernel
    averages come from that sample; the run itself is unchanged)."""
    model = "sme1"
    info = zdm.ModelInfo(model)
    mesh = zdm.Mesh.tri_grid(20, 20, jitter=0.15, seed=6)
    X = state(mesh, info.n_q, 6)
    out = []
    for stride in (0, 1, 3):
        s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
        s.upload(X)
        s.initialize()
        s.enable_timing(stride)
        s.step(9)
        # two flux launches per SSP-RK2 step: the stage's k_flux (id 1) and the final stage
        # fused with the post-step (id 3)
        expect = {0: 0, 1: 9, 3: 3}[stride]
        for kernel in (1, 3):
            ms, n = s.kernel_time(kernel)
            assert n == expect, (stride, kernel, n)
            assert (ms > 0) == (expect > 0)
        assert s.kernel_time(2)[1] == 0
        s.enable_timing(False)
        out.append(s.download()[0])
        s.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
