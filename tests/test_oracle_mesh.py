"""The vectorised SoA mesh builder (product input layer) against the literal
loop transcription of mesh/mesh_reader.py:141-323 on small meshes."""

import numpy as np
import pytest

from oracle import fvm as ofvm, literal


def _pkg():
    from kcfd_amd import mesh as M, workloads as W
    return M, W


def _compare(mv, lm):
    ls = ofvm.soa_from_literal(lm)
    assert mv.n_faces == ls.n_faces and mv.n_cells == ls.n_cells
    assert np.array_equal(mv.owner, ls.owner)
    assert np.array_equal(mv.neighbour, ls.neighbour)
    for a, b in [(mv.face_area, ls.face_area), (mv.face_normal, ls.face_normal),
                 (mv.face_center, ls.face_center), (mv.cell_center, ls.cell_center),
                 (mv.cell_volume, ls.cell_volume)]:
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))
    assert list(mv.boundary_patches) == list(ls.boundary_patches)
    for k in mv.boundary_patches:
        assert np.array_equal(np.asarray(mv.boundary_patches[k]), ls.boundary_patches[k])


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 2), (3, 5), (7, 4)])
def test_quad(nx, ny):
    M, _ = _pkg()
    nodes, cells, bf = M.structured_quad_arrays(0.0, 0.0, 2.0, 1.0, nx, ny)
    mv = M.build_fvmesh_from_arrays(nodes, cells, bf, ndim=2)
    lm = literal.build_mesh(nodes, list(cells), {k: list(v) for k, v in bf.items()}, ndim=2)
    _compare(mv, lm)
    assert mv.n_internal_faces == nx * (ny - 1) + ny * (nx - 1)


@pytest.mark.parametrize("n", [(1, 1, 1), (2, 2, 2), (3, 2, 4)])
def test_hex(n):
    M, _ = _pkg()
    nodes, cells, faces, bf = M.structured_hex_arrays(1.0, 0.5, 2.0, *n)
    mv = M.build_fvmesh_from_arrays(nodes, cells, bf, ndim=3, cell_face_node_list=faces)
    lm = literal.build_mesh(nodes, list(cells), {k: list(v) for k, v in bf.items()}, ndim=3,
                            cell_face_node_list=[list(f) for f in faces])
    _compare(mv, lm)
    # unit-cube hexes have exact volume and area
    h = np.array([1.0 / n[0], 0.5 / n[1], 2.0 / n[2]])
    assert np.allclose(mv.cell_volume, np.prod(h), rtol=1e-14)


def test_tet():
    M, W = _pkg()
    mv = W.tet_mesh(150, seed=2)
    # rebuild the literal mesh from the same tets (recover them from the face lists)
    assert mv.n_cells > 100
    assert np.all(mv.cell_volume > 0)
    # closed surface: sum of outward area vectors of every cell is zero
    o, nb = mv.owner, mv.neighbour
    S = np.zeros((mv.n_cells, 3))
    av = mv.face_normal * mv.face_area[:, None]
    np.add.at(S, o, av)
    np.add.at(S, nb[nb >= 0], -av[nb >= 0])
    assert np.max(np.abs(S)) < 1e-14
    # total volume = unit cube
    assert abs(mv.cell_volume.sum() - 1.0) < 1e-12


def test_tet_literal_small():
    M, _ = _pkg()
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 1, (30, 3))
    tets = Delaunay(pts).simplices
    faces = M.tet_arrays(tets)
    mv = M.build_fvmesh_from_arrays(pts, tets, {}, ndim=3, cell_face_node_list=faces)
    lm = literal.build_mesh(pts, list(tets), {}, ndim=3, cell_face_node_list=[list(f) for f in faces])
    _compare(mv, lm)


def test_structure_counts_config_sizes():
    """Face and nnz counts of the benchmark meshes (closed forms)."""
    from kcfd_amd import device
    _, W = _pkg()
    m = W.quad_mesh(128)
    assert m.n_cells == 16384 and m.n_internal_faces == 32512 and m.n_boundary_faces == 512
    ip, ix, so, sn, ds = device.mesh_structure(m)
    assert ip[-1] == 81408
    m3 = W.hex_mesh(16)
    n = 16
    assert m3.n_internal_faces == 3 * n * n * (n - 1)
    ip, ix, so, sn, ds = device.mesh_structure(m3)
    assert ip[-1] == n ** 3 + 2 * 3 * n * n * (n - 1)
    # the slot map points at the right (row, col) pairs
    rows = np.repeat(np.arange(m3.n_cells), np.diff(ip))
    it = m3.neighbour >= 0
    assert np.array_equal(rows[so[it]], m3.owner[it]) and np.array_equal(ix[so[it]], m3.neighbour[it])
    assert np.array_equal(rows[sn[it]], m3.neighbour[it]) and np.array_equal(ix[sn[it]], m3.owner[it])
    assert np.array_equal(ix[ds], np.arange(m3.n_cells))
