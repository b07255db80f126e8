"""Positivity paths on the device (solver.positivity "mood" / "zhang_shu") against the oracle
on the same seeded wet/dry inputs: bitwise the same states, times and MOOD override counts.
Planted negative-depth cells make the MOOD override fire on every model."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu


def wet_dry_case(model, seed=11):
    info = zdm.ModelInfo(model)
    rng = np.random.default_rng(seed)
    if info.dim == 1:
        mesh = zdm.Mesh.line(96, 0.0, 1.0)
        x = mesh.cell_centroid[:, 0]
        b = np.where(x > 0.6, 0.5, 0.0)
        h = np.maximum(np.where(x < 0.4, 1.0, 0.3) - b, 0.0)
    else:
        mesh = zdm.Mesh.tri_grid(32, 32, jitter=0.15, seed=seed)
        xy = mesh.cell_centroid[:, :2]
        b = np.where(xy[:, 0] > 0.5, 0.5, 0.0) + 0.3 * np.exp(
            -((xy[:, 0] - 0.7) ** 2 + (xy[:, 1] - 0.5) ** 2) / 0.005)
        h = np.maximum(np.where(xy[:, 0] < 0.3, 1.0, 0.4) - b, 0.0)
    planted = rng.choice(np.where(h > 0.2)[0], 6, replace=False)
    h[planted] = -0.5
    X = cases.sme_state(info.n_q, b, h)
    X[:, 2:] = 0.2 * rng.standard_normal((mesh.ncells, info.n_q - 2)) * np.maximum(h, 0)[:, None]
    return info, mesh, X


def run_both(model, positivity, order, steps=12):
    info, mesh, X = wet_dry_case(model)
    print("case", model, positivity, order, flush=True)
    # the MOOD step does not carry RefreshDerivativeAux (as the reference's own MOOD settings,
    # settings.json, turn it off); zhang_shu keeps it on
    refresh = positivity != "mood"
    s = zdm.Solver(model, mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, positivity=positivity,
                   refresh_derivative_aux=refresh)
    A = np.zeros((mesh.ncells, info.n_aux))
    s.upload(X, A)
    s.initialize()
    s.step(steps)
    Xg, Ag = s.download()
    tg, dtg, ng = s.time()
    ntg = s.troubled_cells()
    orc = oracle(model)
    Xc, Ac = X.copy(), A.copy()
    n, tc, dtc = orc.run(mesh, s.face_bc, s.params, Xc, Ac, order=order, ssp_stages=2, cfl=0.45,
                         max_steps=steps, positivity=positivity, refresh_derivative_aux=refresh)
    s.close()
    return (Xg, Ag, tg, dtg, ng, ntg), (Xc, Ac, tc, dtc, n, orc.troubled)


@pytest.mark.parametrize("model", ["sme0_1d", "sme0", "sme1", "sme2", "sme1m"])
@pytest.mark.parametrize("positivity,order", [("mood", 2), ("zhang_shu", 2), ("mood", 1)])
def test_positivity_matches_oracle(model, positivity, order):
    (Xg, Ag, tg, dtg, ng, ntg), (Xc, Ac, tc, dtc, n, ntc) = run_both(model, positivity, order)
    assert ng == n == 12
    assert ntg == ntc
    if positivity == "mood":
        assert ntg > 0
    assert np.array_equal(Xg, Xc), np.max(np.abs(Xg - Xc))
    assert np.array_equal(Ag, Ac), np.max(np.abs(Ag - Ac))
    assert tg == tc and dtg == dtc


def test_mood_ensemble_matches_single_runs():
    model = "sme1"
    info, mesh, X = wet_dry_case(model)
    info2, mesh2, X2 = wet_dry_case(model, seed=12)
    u, offs = zdm.Mesh.concat([mesh, mesh2])
    e = zdm.Solver(model, u, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, positivity="mood",
                   cell_offsets=offs)
    e.upload(np.vstack([X, X2]))
    e.initialize()
    e.step(8)
    Xe, _ = e.download()
    nte = e.troubled_cells()
    e.close()
    parts, nts = [], 0
    for m, Xm in ((mesh, X), (mesh2, X2)):
        s = zdm.Solver(model, m, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, positivity="mood")
        s.upload(Xm)
        s.initialize()
        s.step(8)
        parts.append(s.download()[0])
        nts += s.troubled_cells()
        s.close()
    assert nte == nts > 0
    assert np.array_equal(Xe, np.vstack(parts))


def test_positivity_step_graph_replays_match_eager():
    model = "sme2"
    info, mesh, X = wet_dry_case(model)
    out = []
    for graph in (False, True):
        s = zdm.Solver(model, mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, positivity="mood")
        s.upload(X)
        s.initialize()
        for _ in range(3):
            s.step_graph(4) if graph else s.step(4)
            s.synchronize()
        out.append((s.download()[0], s.troubled_cells()))
        s.close()
    assert out[0][1] == out[1][1] > 0
    assert np.array_equal(out[0][0], out[1][0])
