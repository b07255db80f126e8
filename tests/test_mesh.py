"""Mesh service (host): connectivity, geometry and tags against closed forms."""
import os

import numpy as np
import pytest

from conftest import ROOT, zdm

GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("renumber", ["rcm", "hilbert", "none"])
@pytest.mark.parametrize("n,jitter", [(4, 0.0), (7, 0.15), (16, 0.0)])
def test_tri_grid_counts_and_geometry(n, jitter, renumber):
    m = zdm.Mesh.tri_grid(n, n, jitter=jitter, seed=9, renumber=renumber)
    # an n x n split grid has 2n^2 cells, 3n^2 + 2n faces, 4n boundary faces
    assert m.ncells == 2 * n * n
    assert m.nfaces == 3 * n * n + 2 * n
    bnd = m.face_cells[:, 1] < 0
    assert bnd.sum() == 4 * n
    assert np.all(m.cell_volume > 0)
    assert abs(m.cell_volume.sum() - 1.0) <= 1e-14
    # every boundary face carries tag 1 (no physical names -> DMPlexMarkBoundaryFaces(1))
    assert np.all(m.face_tag[bnd] == 1) and np.all(m.face_tag[~bnd] == -1)
    # outward normals of each cell sum to zero (closed polygon)
    acc = np.zeros((m.ncells, 3))
    for f in range(m.nfaces):
        L, R = m.face_cells[f]
        acc[L] += m.face_normal[f]
        if R >= 0:
            acc[R] -= m.face_normal[f]
    assert np.max(np.abs(acc)) <= 1e-15
    # orientation support[0] -> support[1]; outward on the boundary
    L, R = m.face_cells[:, 0], m.face_cells[:, 1]
    inter = R >= 0
    d = m.cell_centroid[R[inter]] - m.cell_centroid[L[inter]]
    assert np.all(np.sum(d * m.face_normal[inter], axis=1) > 0)
    d = m.face_centroid[bnd] - m.cell_centroid[L[bnd]]
    assert np.all(np.sum(d * m.face_normal[bnd], axis=1) > 0)
    assert np.all(L[inter] < R[inter])
    # faces are ordered by (support[0], support[1]) and every cell lists its faces ascending
    assert np.all(np.diff(L) >= 0)
    for c in range(m.ncells):
        fs = m.cell_faces[c, :m.cell_nfaces[c]]
        assert np.all(np.diff(fs) > 0)
        for k, f in enumerate(fs):
            a, b = m.face_cells[f]
            assert c in (a, b)
            assert m.cell_nbrs[c, k] == (b if a == c else a)
    assert sorted(m.cell_perm.tolist()) == list(range(m.ncells))


def test_triangle_geometry_closed_form():
    m = zdm.Mesh.tri_grid(1, 1, renumber=False)
    # lower-right (0,0),(1,0),(1,1) and upper-left (0,0),(1,1),(0,1)
    c = {tuple(np.round(x[:2], 12)) for x in m.cell_centroid}
    assert c == {(round(2 / 3, 12), round(1 / 3, 12)), (round(1 / 3, 12), round(2 / 3, 12))}
    assert np.allclose(m.cell_volume, 0.5)
    # the diagonal face: normal of length sqrt(2), midpoint (1/2, 1/2)
    inter = np.nonzero(m.face_cells[:, 1] >= 0)[0]
    assert len(inter) == 1
    f = inter[0]
    assert np.allclose(np.linalg.norm(m.face_normal[f]), np.sqrt(2))
    assert np.allclose(m.face_centroid[f, :2], [0.5, 0.5])


def test_rcm_reduces_bandwidth():
    m = zdm.Mesh.tri_grid(40, 40, jitter=0.1, seed=3)
    inter = m.face_cells[:, 1] >= 0
    bw = np.max(m.face_cells[inter, 1] - m.face_cells[inter, 0])
    assert bw <= 3 * 40 + 4


def test_hilbert_order_locality():
    """Hilbert order keeps face neighbours close: most interior faces join cells less than a
    wavefront (64 cells) apart."""
    m = zdm.Mesh.tri_grid(64, 64, renumber="hilbert")
    inter = m.face_cells[:, 1] >= 0
    gap = m.face_cells[inter, 1] - m.face_cells[inter, 0]
    assert np.mean(gap < 64) > 0.9


def test_line_mesh():
    m = zdm.Mesh.line(10, 0.0, 2.0)
    assert m.dim == 1 and m.ncells == 10 and m.nfaces == 11
    assert np.allclose(m.cell_volume, 0.2)
    assert np.all(np.abs(m.face_normal[:, 0]) == 1.0)
    bnd = np.nonzero(m.face_cells[:, 1] < 0)[0]
    assert sorted(m.face_tag[bnd].tolist()) == [1, 2]


def test_reference_1d_mesh_fixture_and_tag_map():
    """mesh_n60.msh with the reference's 1-D SME-0 model: only "default" is a model tag, the
    mesh has physical names, so "default" maps tag 1 (the "left" point) and the "right" point
    (tag 2) gets no boundary flux (VirtualSolver.hpp:500-511, TransportStep.hpp:498)."""
    m = zdm.Mesh.read_msh(os.path.join(GOLDEN, "mesh_n60.msh"))
    assert m.dim == 1 and m.ncells == 60 and m.nfaces == 61
    names = {n: (d, t) for n, d, t in m.physical}
    assert names["left"] == (0, 1) and names["right"] == (0, 2) and names["domain"] == (1, 3)
    assert abs(m.cell_volume.sum() - 3.0) <= 1e-12
    fb = zdm.boundary_face_bc(m, ["default"], 1)
    bnd = np.nonzero(m.face_cells[:, 1] < 0)[0]
    xs = {round(float(m.face_centroid[f, 0]), 6): int(fb[f]) for f in bnd}
    assert xs == {-1.5: 0, 1.5: -1}


def test_reference_quad_strip_and_rotation():
    a = zdm.Mesh.read_msh(os.path.join(GOLDEN, "bump_2d.msh"))
    b = zdm.Mesh.read_msh(os.path.join(GOLDEN, "bump_2d_rot.msh"))
    for m in (a, b):
        assert m.dim == 2 and m.ncells == 960 and m.max_faces == 4
        assert np.all(m.cell_volume > 0)
        assert np.sum(m.face_cells[:, 1] < 0) == 256
        tags = sorted(set(m.face_tag[m.face_cells[:, 1] < 0].tolist()))
        assert tags == [1, 2, 3, 4]
    assert abs(a.cell_volume.sum() - b.cell_volume.sum()) <= 1e-12
    # a model tag missing from the mesh is refused (VirtualSolver.hpp:407-417)
    with pytest.raises(zdm.ZdmError):
        zdm.boundary_face_bc(a, ["inflow"], 2)
    bnd = a.face_cells[:, 1] < 0
    fb = zdm.boundary_face_bc(a, ["left", "right"], 2)
    assert np.all(fb[bnd & (a.face_tag == 1)] == 0)
    assert np.all(fb[bnd & (a.face_tag == 2)] == 1)
    assert np.all(fb[bnd & (a.face_tag > 2)] == -1)
    assert np.all(fb[~bnd] == -1)
    # "default" maps tag 1 as well; the later model index wins (VirtualSolver.hpp:510)
    fb = zdm.boundary_face_bc(a, ["left", "default"], 2)
    assert np.all(fb[bnd & (a.face_tag == 1)] == 1)


def test_splitmix64_known_values():
    import ctypes
    # SplitMix64 (Vigna): the first output of seed 0 is 0xE220A8397B1DCDAF
    st = ctypes.c_uint64(0)
    v = zdm.mesh_lib().zdm_splitmix64_next(ctypes.byref(st))
    assert v == 0xE220A8397B1DCDAF
    from conftest import cases
    r = cases.SplitMix64(0)
    assert r.next() == 0xE220A8397B1DCDAF
    r2 = cases.SplitMix64(12345)
    st = ctypes.c_uint64(12345)
    for _ in range(5):
        assert r2.next() == zdm.mesh_lib().zdm_splitmix64_next(ctypes.byref(st))


def test_mesh_concat_is_a_disjoint_union():
    ms = [zdm.Mesh.tri_grid(n, n, jitter=0.1, seed=n) for n in (3, 6, 4)]
    u, off = zdm.Mesh.concat(ms)
    assert list(off) == [0, 18, 90, 122] and u.ncells == 122
    assert u.nfaces == sum(m.nfaces for m in ms)
    fo = 0
    for i, m in enumerate(ms):
        sl = slice(off[i], off[i + 1])
        assert np.array_equal(u.cell_volume[sl], m.cell_volume)
        assert np.array_equal(u.cell_nbrs[sl], np.where(m.cell_nbrs >= 0, m.cell_nbrs + off[i], -1))
        assert np.array_equal(u.cell_faces[sl, :m.max_faces],
                              np.where(m.cell_faces >= 0, m.cell_faces + fo, -1))
        fo += m.nfaces
    # faces stay ordered by (support[0], support[1]) and every neighbour is in its member
    assert np.all(np.diff(u.face_cells[:, 0]) >= 0)
    member = np.searchsorted(off, np.arange(u.ncells), side="right") - 1
    inner = u.face_cells[:, 1] >= 0
    assert np.array_equal(member[u.face_cells[inner, 0]], member[u.face_cells[inner, 1]])
