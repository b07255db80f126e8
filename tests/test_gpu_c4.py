"""The ensemble on its full-size workload: 256 SME level-1 members (cases.ensemble_member:
jittered grids of 10 k - 200 k cells, 8-bump beds, dam breaks at x0 ~ U(0.3, 0.7) with depth
ratios U(0.3, 0.8); 17.7 M cells in all), MUSCL with the limiter, implicit slip source,
SSP-RK2 at CFL 0.45, graph-replayed batched steps at the autotuned launch shape.

Every member's water volume is conserved to round-off and its state stays finite; the members
take their own time steps; and sampled members (the smallest, the largest and two seeded
others) are bitwise equal to the same runs on their own solvers, with the same t, dt and step
count.  The small-ensemble tests compare members with the oracle."""
import numpy as np
import pytest

from conftest import cases, zdm

pytestmark = pytest.mark.gpu

COUNT = 256
STEPS = 4
KW = dict(order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True, use_limiters=True)


def test_c4_full_size_batched_members_equal_single_runs():
    meshes, states = [], []
    for i in range(COUNT):
        n, mesh_seed, state = cases.ensemble_member(i)
        m = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=mesh_seed)
        meshes.append(m)
        states.append(state(m))
    sizes = [m.ncells for m in meshes]
    print("c4: %d members, %d - %d cells, %d in all" % (COUNT, min(sizes), max(sizes), sum(sizes)), flush=True)
    e = zdm.Ensemble("sme1", meshes, **KW)
    e.upload(e.join(states))
    e.initialize()
    bs, cpt, _ = e.autotune(3)
    print("  autotuned launch %d x %d" % (bs, cpt), flush=True)
    e.step_graph(STEPS)
    Xe, _ = e.download()
    times = [e.member_time(i) for i in range(COUNT)]
    e.close()
    assert np.all(np.isfinite(Xe))
    Xm = e.split(Xe)
    for i, (m, X0, X1) in enumerate(zip(meshes, states, Xm)):
        v0 = float(np.sum(m.cell_volume * X0[:, 1]))
        v1 = float(np.sum(m.cell_volume * X1[:, 1]))
        assert abs(v1 - v0) <= 1e-13 * v0, (i, v0, v1)
        assert np.array_equal(X1[:, 0], X0[:, 0]), i
        assert times[i][2] == STEPS, (i, times[i])
    assert len({dt for _, dt, _ in times}) > COUNT // 2
    rng = np.random.default_rng(4)
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} | set(rng.choice(COUNT, 2, replace=False).tolist()))
    for i in sample:
        print("  single run of member %d (%d cells)" % (i, sizes[i]), flush=True)
        s = zdm.Solver("sme1", meshes[i], **KW)
        s.upload(states[i])
        s.initialize()
        s.step_graph(STEPS)
        Xs, _ = s.download()
        ts = s.time()
        s.close()
        assert np.array_equal(Xm[i], Xs), i
        assert ts == times[i], (i, ts, times[i])
