"""Device Chorin pressure split (VAM) against the CPU oracle.

Bitwise: derivative aux, pressure residual and the assembled Pmat use the same correctly
rounded operations in the same per-cell order as the oracle.  Krylov results differ from the
oracle only through the summation order of the dot products, so the solved pressure and the
state after a step are compared to the Krylov tolerance (rtol times the conditioning)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import ROOT, cases, oracle, zdm
from test_oracle_vam import c3_small, read_petsc_mat

pytestmark = pytest.mark.gpu


def vam_case(n=14, seed=33):
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=seed, renumber=True)
    return mesh


def bsr_to_csr(rp, col, val, nb):
    rows, cols, vals = [], [], []
    for c in range(nb):
        for k in range(rp[c], rp[c + 1]):
            for a in range(2):
                for b in range(2):
                    rows.append(2 * c + a)
                    cols.append(2 * col[k] + b)
                    vals.append(val[k, a, b])
    return sp.csr_matrix((vals, (rows, cols)), shape=(2 * nb, 2 * nb))


def csr_to_bsr(A):
    nb = A.shape[0] // 2
    D = A.tocsr()
    rp, col, val = [0], [], []
    for c in range(nb):
        blocks = sorted(set(D.indices[D.indptr[2 * c]:D.indptr[2 * c + 2]] // 2))
        for bcol in blocks:
            col.append(bcol)
            val.append(D[2 * c:2 * c + 2, 2 * bcol:2 * bcol + 2].toarray())
        rp.append(len(col))
    return np.array(rp, dtype=np.int32), np.array(col, dtype=np.int32), np.array(val)


def test_frozen_aux_residual_pmat_bitwise():
    mesh = vam_case()
    print("case chorin bitwise cells", mesh.ncells, flush=True)
    orc = oracle("vam")
    X, CA_ref = c3_small(mesh, orc)
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9)
    s.upload(X)
    CA = s.vam_frozen_aux()
    frozen = [j for j in range(20) if j >= 8]
    assert np.array_equal(CA[:, frozen], CA_ref[:, frozen])
    dt = 2e-3
    rng = np.random.default_rng(4)
    for scale in (0.0, 1.0, 30.0):
        P = scale * rng.standard_normal((mesh.ncells, 2))
        R = s.vam_residual(dt, P)
        R_ref = orc.pressure_residual(mesh, s.params, dt, X, P, CA_ref.copy())
        assert np.array_equal(R, R_ref)
    rp, col, val = s.vam_pmat(dt)
    Pm = bsr_to_csr(rp, col, val, mesh.ncells)
    Pm_ref = orc.assemble_pmat(mesh, s.params, dt, X, CA_ref)
    assert (Pm != Pm_ref).nnz == 0
    # the stencil pattern: the cell itself and its two-ring, ascending block columns
    for c in range(mesh.ncells):
        cc = col[rp[c]:rp[c + 1]]
        assert c in cc and np.all(np.diff(cc) > 0)
    s.close()


def test_pmat_fixture_gmres_device():
    A = read_petsc_mat(os.path.join(ROOT, "tests", "golden", "pmat.bin"))
    rp, col, val = csr_to_bsr(A)
    assert np.array_equal(bsr_to_csr(rp, col, val, 60).toarray(), A.toarray())
    rng = np.random.default_rng(11)
    b = rng.standard_normal(120)
    print("case pmat.bin gmres", flush=True)
    x, its, reason = zdm.bsr2_gmres(rp, col, val, b, restart=30, max_it=400, rtol=1e-10, atol=1e-50)
    assert reason == 2
    xd = spla.spsolve(A.tocsc(), b)
    cond = np.linalg.cond(A.toarray())
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 10 * cond * 1e-10
    xo, reason_o, its_o = oracle("vam").csr_gmres(A, b, restart=30, max_it=400, rtol=1e-10)
    assert reason_o == 2 and abs(its - its_o) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 10 * cond * 1e-10
    # a few iterations only: the same Krylov iterate as the oracle up to round-off
    for k in (1, 5, 20):
        x, its, reason = zdm.bsr2_gmres(rp, col, val, b, restart=30, max_it=k, rtol=1e-30)
        xo, _, _ = oracle("vam").csr_gmres(A, b, restart=30, max_it=k, rtol=1e-30)
        assert its == k and reason == -3
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("op_mode", [0, 1])
def test_chorin_steps_match_oracle(op_mode):
    mesh = zdm.Mesh.tri_grid(10, 10, jitter=0.15, seed=3, renumber=True)
    print("case chorin steps op_mode", op_mode, "cells", mesh.ncells, flush=True)
    orc = oracle("vam")
    X = cases.c3_state(mesh)
    A = np.zeros((mesh.ncells, orc.n_aux))
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9)
    kw = dict(restart=30, max_it=400, rtol=1e-10, atol=1e-50)
    s.chorin_setup(op_mode=op_mode, **kw)
    s.upload(X, A)
    s.initialize()
    # step 1: the predictor (and so h) is bitwise; P and the corrected moments to tolerance
    n, its, reason = s.chorin_step(1)
    Xc, Ac, Pc = X.copy(), A.copy(), np.zeros((mesh.ncells, 2))
    nc, tc, its_c, reason_c = orc.chorin_run(mesh, s.face_bc, s.params, Xc, Ac, Pc, order=1,
                                             max_steps=1, op_mode=op_mode, **kw)
    Xg, _ = s.download()
    Pg = s.chorin_pressure()
    assert n == nc == 1 and reason == reason_c
    assert abs(its - its_c) <= 2
    assert np.array_equal(Xg[:, :2], Xc[:, :2])
    tol = 1e-7
    assert np.linalg.norm(Pg - Pc) <= tol * np.linalg.norm(Pc)
    assert np.max(np.abs(Xg - Xc)) <= tol * np.max(np.abs(Xc))
    # more steps: mass to round-off on the device, state to the Krylov tolerance
    mass0 = np.dot(X[:, 1], mesh.cell_volume)
    n, its, reason = s.chorin_step(4)
    nc, tc, its_c, reason_c = orc.chorin_run(mesh, s.face_bc, s.params, Xc, Ac, Pc, order=1,
                                             max_steps=4, op_mode=op_mode, time=tc, **kw)
    Xg, _ = s.download()
    tg, _, steps = s.time()
    assert n == nc == 4 and steps == 5
    assert abs(tg - tc) <= 1e-12 * tc
    assert abs(np.dot(Xg[:, 1], mesh.cell_volume) - mass0) <= 1e-13 * mass0
    assert np.max(np.abs(Xg - Xc)) <= 1e-6 * np.max(np.abs(Xc))
    s.close()


def test_chorin_operator_modes_agree():
    """Matrix-free Aop and the assembled Pmat solve the same system (ChorinVAM.hpp:382)."""
    mesh = zdm.Mesh.tri_grid(12, 12, jitter=0.15, seed=5, renumber=True)
    X = cases.c3_state(mesh)
    out = []
    for op_mode in (0, 1):
        s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9)
        s.chorin_setup(restart=30, max_it=400, rtol=1e-10, atol=1e-50, op_mode=op_mode)
        s.upload(X)
        s.initialize()
        s.chorin_step(1)
        out.append(s.chorin_pressure())
        s.close()
    assert np.linalg.norm(out[0] - out[1]) <= 1e-6 * np.linalg.norm(out[1])


def test_chorin_needs_vam_model():
    mesh = zdm.Mesh.tri_grid(4, 4)
    s = zdm.Solver("sme1", mesh, order=1)
    with pytest.raises(zdm.ZdmError, match="VAM"):
        s.chorin_step(1)
    s.close()
