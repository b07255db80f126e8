"""zdm_transport_rhs_device: the transport right-hand side on device buffers, ordered on a
caller's stream, against the host-array entry point (bitwise) and the oracle."""
import numpy as np
import pytest

from conftest import cases, oracle, zdm

pytestmark = pytest.mark.gpu


def _case(model, n=14, seed=31):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed)
    info = zdm.ModelInfo(model)
    X = cases.c1_state(mesh, info.n_q, seed=seed)
    rng = np.random.default_rng(seed)
    X[:, 2:] = 0.05 * rng.standard_normal((mesh.ncells, info.n_q - 2))
    A = np.zeros((mesh.ncells, info.n_aux))
    return mesh, X, A


@pytest.mark.parametrize("model,order", [("sme2", 2), ("sme1", 2), ("sme0", 1)])
def test_device_rhs_on_a_side_stream_matches_host_entry_point(model, order):
    torch = pytest.importorskip("torch")
    mesh, X, A = _case(model)
    s = zdm.Solver(model, mesh, order=order, ssp_stages=2, cfl=0.45, t_end=1e9, use_limiters=True)
    print("case", model, "order", order, "cells", mesh.ncells, flush=True)
    F_host = s.transport_rhs(X, A, t=0.25)
    side = torch.cuda.Stream()
    base = torch.tensor(X, dtype=torch.float64, device="cuda")
    Ad = torch.tensor(A, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        Xd = base * 1.0                      # produced on the side stream
        Fd = torch.full_like(Xd, float("nan"))
        s.transport_rhs_device(Xd.data_ptr(), Fd.data_ptr(), Ad.data_ptr(), t=0.25,
                               stream=side.cuda_stream)
        G = Fd * 1.0                         # consumed on the side stream
    side.synchronize()
    assert np.array_equal(G.cpu().numpy(), F_host)
    # without an aux input, on the solver's own stream
    F_host0 = s.transport_rhs(X, None, t=0.0)
    F0 = torch.full_like(base, float("nan"))
    torch.cuda.synchronize()
    s.transport_rhs_device(base.data_ptr(), F0.data_ptr(), None, t=0.0)
    s.synchronize()
    assert np.array_equal(F0.cpu().numpy(), F_host0)
    # and the host entry point itself against the oracle's FormRHS
    Fo = oracle(model).form_rhs(mesh, s.face_bc, s.params, X.copy(), A.copy(), order=order, t=0.25)
    assert np.max(np.abs(F_host - Fo)) <= 1e-12 * max(np.max(np.abs(Fo)), 1e-300)
    s.close()
