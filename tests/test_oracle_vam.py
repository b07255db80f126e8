"""Pins the oracle's Chorin VAM pressure split (CPU).

  * the reference's dumped pressure matrix tests/golden/pmat.bin: GMRES(30) + point-block
    Jacobi against a direct solve, and the preconditioned-residual stopping test;
  * GMRES as a minimum-residual method: after k iterations of the first cycle the iterate
    minimises |M^-1 (b - A x)| over the Krylov space, computed here by numpy least squares;
  * Green-Gauss derivative: exact for linear fields on interior cells (closed form), and the
    zero-Neumann boundary rule on a two-cell brute force;
  * the reference's own consistency gate (ChorinVAM.hpp:323-327, 382): the assembled Pmat
    applied to P equals residual(P) - residual(0) to 1e-10, which compares the two separate
    transcriptions vam_pressure_blocks / vam_pressure_residual;
  * a Chorin step runs and conserves the water column (the corrector leaves b and h alone).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import ROOT, cases, oracle, zdm

GOLD = os.path.join(ROOT, "tests", "golden")


def read_petsc_mat(path):
    """PETSc binary MatView: big-endian int32 [classid, m, n, nnz], row lengths, columns,
    float64 values."""
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype=">i4")
    assert hdr[0] == 1211216
    m, n, nnz = int(hdr[1]), int(hdr[2]), int(hdr[3])
    off = 16
    rl = np.frombuffer(raw[off:off + 4 * m], dtype=">i4").astype(np.int64)
    off += 4 * m
    ci = np.frombuffer(raw[off:off + 4 * nnz], dtype=">i4").astype(np.int32)
    off += 4 * nnz
    v = np.frombuffer(raw[off:off + 8 * nnz], dtype=">f8").astype(np.float64)
    off += 8 * nnz
    assert off == len(raw) and rl.sum() == nnz
    rp = np.concatenate([[0], np.cumsum(rl)]).astype(np.int32)
    return sp.csr_matrix((v, ci, rp), shape=(m, n))


@pytest.fixture(scope="module")
def vam():
    return oracle("vam")


@pytest.fixture(scope="module")
def pmat():
    A = read_petsc_mat(os.path.join(GOLD, "pmat.bin"))
    assert A.shape == (120, 120) and A.nnz == 1060
    assert open(os.path.join(GOLD, "pmat.bin.info")).read().split() == ["-matload_block_size", "2"]
    return A


def pbj_inverse(A):
    n = A.shape[0]
    D = np.zeros((n // 2, 2, 2))
    Ad = A.toarray() if n <= 4000 else None
    for c in range(n // 2):
        D[c] = np.linalg.inv(Ad[2 * c:2 * c + 2, 2 * c:2 * c + 2])
    return sp.block_diag(list(D), format="csr")


def test_pmat_fixture_structure(pmat):
    """Block size 2: every row pair shares its column blocks; asymmetric operator."""
    A = pmat
    asym = spla.norm(A - A.T) / spla.norm(A)
    assert 1.2 < asym < 1.35
    for c in range(60):
        r0 = set(A.indices[A.indptr[2 * c]:A.indptr[2 * c + 1]] // 2)
        r1 = set(A.indices[A.indptr[2 * c + 1]:A.indptr[2 * c + 2]] // 2)
        assert r0 == r1 and c in r0


def test_pmat_gmres_pbj_vs_direct(vam, pmat):
    rng = np.random.default_rng(11)
    b = rng.standard_normal(120)
    x, reason, its = vam.csr_gmres(pmat, b, restart=30, max_it=400, rtol=1e-10)
    assert reason == 2 and 1 <= its < 400
    xd = spla.spsolve(pmat.tocsc(), b)
    cond = np.linalg.cond(pmat.toarray())
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 10 * cond * 1e-10
    # KSPConvergedDefault on the preconditioned residual norm (left preconditioning)
    Minv = pbj_inverse(pmat)
    assert np.linalg.norm(Minv @ (b - pmat @ x)) <= 1e-10 * np.linalg.norm(Minv @ b) * (1 + 1e-12)


def test_gmres_first_cycle_is_minimum_residual(vam, pmat):
    """k < restart iterations from x0 = 0: x_k = argmin |M^-1(b - A x)| over
    span{z, Bz, ..., B^(k-1) z}, z = M^-1 b, B = M^-1 A (left-preconditioned GMRES)."""
    rng = np.random.default_rng(5)
    b = rng.standard_normal(120)
    Minv = pbj_inverse(pmat).toarray()
    B = Minv @ pmat.toarray()
    z = Minv @ b
    for k in (1, 3, 7, 12):
        x, reason, its = vam.csr_gmres(pmat, b, restart=30, max_it=k, rtol=1e-30)
        assert reason == -3 and its == k
        K = np.empty((120, k))
        v = z / np.linalg.norm(z)
        for j in range(k):          # orthonormal Krylov basis (modified Gram-Schmidt, twice)
            w = v if j == 0 else B @ K[:, j - 1]
            for _ in range(2):
                w = w - K[:, :j] @ (K[:, :j].T @ w)
            K[:, j] = w / np.linalg.norm(w)
        y, *_ = np.linalg.lstsq(B @ K, z, rcond=None)
        xr = K @ y
        assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)


def test_gmres_restart_and_exact_small(vam):
    """A nonsymmetric 2x2-block system of 8 unknowns: GMRES(8) is exact in <= 8 iterations;
    GMRES(3) with restarts reaches the same answer."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((8, 8)) + 6 * np.eye(8)
    As = sp.csr_matrix(A)
    b = rng.standard_normal(8)
    xd = np.linalg.solve(A, b)
    for restart in (8, 3):
        x, reason, its = vam.csr_gmres(As, b, restart=restart, max_it=200, rtol=1e-13)
        assert reason > 0
        assert np.allclose(x, xd, rtol=1e-10, atol=1e-12)


@pytest.fixture(scope="module")
def jmesh():
    return zdm.Mesh.tri_grid(14, 14, jitter=0.15, seed=33, renumber=True)


def test_deriv1_linear_field_exact_interior(vam, jmesh):
    """On the unjittered split grid the mean of two adjacent centroids is the face midpoint,
    so the mean-face Green-Gauss rule is exact for linear fields on interior cells."""
    m = zdm.Mesh.tri_grid(12, 12, renumber=True)
    xy = m.cell_centroid[:, :2]
    f = 0.7 + 1.3 * xy[:, 0] - 2.1 * xy[:, 1]
    interior = (m.cell_nbrs >= 0).sum(axis=1) == m.cell_nfaces
    for axis, slope in ((0, 1.3), (1, -2.1)):
        d = vam.deriv1(m, f, axis)
        assert np.max(np.abs(d[interior] - slope)) <= 1e-11
        # a constant field has zero derivative everywhere (closed surface: sum of N = 0)
        d0 = vam.deriv1(jmesh, np.full(jmesh.ncells, 3.0), axis)
        assert np.max(np.abs(d0)) <= 1e-10


def test_deriv1_two_cell_brute_force(vam):
    """Two triangles of the unit square: interior face value = mean, boundary faces take the
    cell value (zero Neumann, UserFunctions.H:99-103), then / volume."""
    m = zdm.Mesh.tri_grid(1, 1, renumber=False)
    f = np.array([2.0, 5.0])
    for axis in (0, 1):
        d = vam.deriv1(m, f, axis)
        ref = np.zeros(2)
        for fi in range(m.nfaces):
            L, R = m.face_cells[fi]
            N = m.face_normal[fi, axis]
            if R >= 0:
                ref[L] += 0.5 * (f[L] + f[R]) * N
                ref[R] -= 0.5 * (f[L] + f[R]) * N
            else:
                ref[L] += f[L] * N
        ref /= m.cell_volume
        assert np.allclose(d, ref, rtol=0, atol=1e-15)
        # closed form: sum_f N = 0 on each cell leaves (fR - fL)/2 N_LR / V, the same on both
        fi = int(np.nonzero(m.face_cells[:, 1] >= 0)[0][0])
        L, R = m.face_cells[fi]
        want = (f[R] - f[L]) / 2 * m.face_normal[fi, axis] / 0.5
        assert abs(abs(want) - 3.0) <= 1e-14
        assert np.allclose(d, [want, want], rtol=0, atol=1e-14)


def c3_small(mesh, vam):
    X = cases.c3_state(mesh)
    # give the moments some structure so every term of the operator is exercised
    xy = mesh.cell_centroid[:, :2]
    X[:, 3] = 0.05 * np.sin(3 * xy[:, 0])
    X[:, 4] = 0.1 * np.cos(2 * xy[:, 1])
    X[:, 5] = 0.02 * xy[:, 0] * xy[:, 1]
    X[:, 6] = 0.01 * np.sin(xy[:, 1])
    X[:, 7] = -0.01 * xy[:, 0]
    CA = np.zeros((mesh.ncells, 20))
    vam.fill_ca(mesh, False, X, np.zeros((mesh.ncells, 2)), CA)
    return X, CA


def test_affine_consistency_gate(vam, jmesh):
    """ChorinVAM.hpp:323-327, 382: |Pmat P - (residual(P) - residual(0))| / |.| < 1e-10."""
    m = jmesh
    p = vam.default_params
    X, CA = c3_small(m, vam)
    dt = 2e-3
    Pm = vam.assemble_pmat(m, p, dt, X, CA)
    rng = np.random.default_rng(9)
    R0 = vam.pressure_residual(m, p, dt, X, np.zeros((m.ncells, 2)), CA.copy())
    for _ in range(3):
        P = rng.standard_normal((m.ncells, 2))
        R = vam.pressure_residual(m, p, dt, X, P, CA.copy())
        lhs = Pm @ P.ravel()
        rhs = (R - R0).ravel()
        assert np.linalg.norm(lhs - rhs) <= 1e-12 * np.linalg.norm(rhs)
    # block structure: 2x2 blocks over the two-ring stencil (<= 10 block columns on triangles)
    bcols = [len(set(Pm.indices[Pm.indptr[2 * c]:Pm.indptr[2 * c + 2]] // 2)) for c in range(m.ncells)]
    assert max(bcols) <= 13 and min(bcols) >= 1


def test_residual_is_affine(vam, jmesh):
    m = jmesh
    p = vam.default_params
    X, CA = c3_small(m, vam)
    rng = np.random.default_rng(2)
    P1, P2 = rng.standard_normal((2, m.ncells, 2))
    r = lambda P: vam.pressure_residual(m, p, 1e-3, X, P, CA.copy())
    lhs = r(0.3 * P1 + 0.7 * P2)
    rhs = 0.3 * r(P1) + 0.7 * r(P2)
    assert np.linalg.norm(lhs - rhs) <= 1e-12 * np.linalg.norm(rhs)


@pytest.mark.parametrize("op_mode", [0, 1])
def test_chorin_step_runs(vam, op_mode):
    m = zdm.Mesh.tri_grid(10, 10, jitter=0.15, seed=3, renumber=True)
    p = vam.default_params
    X = cases.c3_state(m)
    A = np.zeros((m.ncells, vam.n_aux))
    P = np.zeros((m.ncells, 2))
    bc = zdm.boundary_face_bc(m, ["default"], 2)
    mass0 = np.dot(X[:, 1], m.cell_volume)
    n, t, its, reason = vam.chorin_run(m, bc, p, X, A, P, order=1, max_steps=2, op_mode=op_mode)
    assert n == 2 and t > 0 and its >= 1
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(P))
    # walls: h is advanced by the conservative predictor only (the corrector leaves b, h)
    assert abs(np.dot(X[:, 1], m.cell_volume) - mass0) <= 1e-13 * mass0
