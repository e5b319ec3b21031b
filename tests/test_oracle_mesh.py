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


HYBRID_ARGS = [(2.0, 0.1, 0.1, 4, 2, 2, 0.5), (1.0, 0.5, 0.3, 5, 3, 2, 0.4),
               (2.0, 0.1, 0.1, 3, 1, 1, 0.0), (1.0, 1.0, 1.0, 2, 2, 2, 1.0)]


@pytest.mark.parametrize("args", HYBRID_ARGS)
def test_hybrid_generator_equals_literal(args):
    """The vectorised hybrid hex/tet generator against the loop restatement of
    mesh/hybrid_mesh_generator.py:93-270: identical nodes, cells, per-cell
    face lists, boundary face lists (in order) and zones; the built meshes
    equal as for the uniform meshes."""
    M, _ = _pkg()
    nodes, rag, bd, zones = M.hybrid_hex_tet_arrays(*args)
    ln, lc, lf, lbd, lz = literal.hybrid_hex_tet(*args)
    assert np.array_equal(nodes, ln)
    assert rag.cell_nodes.shape[0] == len(lc)
    for c in range(len(lc)):
        k = len(lc[c])
        assert np.array_equal(rag.cell_nodes[c, :k], lc[c]) and np.all(rag.cell_nodes[c, k:] == -1)
        seg = rag.faces[rag.face_ptr[c]:rag.face_ptr[c + 1]]
        assert len(seg) == len(lf[c])
        for f, r in zip(lf[c], seg):
            assert np.array_equal(r[:len(f)], f) and np.all(r[len(f):] == -1)
    assert list(bd) == list(lbd)
    for nm in bd:
        assert len(bd[nm]) == len(lbd[nm])
        for r, f in zip(bd[nm], lbd[nm]):
            assert np.array_equal(r[:len(f)], f) and np.all(r[len(f):] == -1)
    assert list(zones["hex"]) == lz["hex"] and list(zones["tet"]) == lz["tet"]
    mv = M.build_fvmesh_from_arrays(nodes, None, bd, cell_zone_map=zones, ndim=3,
                                    cell_face_node_list=rag)
    _compare(mv, literal.build_mesh(ln, lc, lbd, ndim=3, cell_face_node_list=lf))
    # the reference's own list-of-lists input takes the same ragged path
    mv2 = M.build_fvmesh_from_arrays(ln, lc, lbd, ndim=3, cell_face_node_list=lf)
    assert np.array_equal(mv2.owner, mv.owner) and np.array_equal(mv2.cell_volume, mv.cell_volume)


def test_hybrid_closed_forms():
    """Counts, volumes (hex dx dy dz; each of a hex's 12 centre tets
    dx dy dz / 12), outer areas per patch, closed cells, and a conforming
    hex/tet interface (every interface triangle internal, no 'default'
    patch)."""
    M, _ = _pkg()
    Lx, Ly, Lz, nx, ny, nz, tf = 2.0, 0.1, 0.1, 8, 3, 2, 0.5
    m = M.generate_hybrid_hex_tet_mesh(Lx, Ly, Lz, nx, ny, nz, tf)
    nxh = max(1, int(nx * (1 - tf)))
    nxt = nx - nxh
    nhex, ntet = nxh * ny * nz, 12 * nxt * ny * nz
    assert m.n_cells == nhex + ntet
    assert len(m.cell_zones["hex"]) == nhex and len(m.cell_zones["tet"]) == ntet
    h = (Lx / nx) * (Ly / ny) * (Lz / nz)
    assert np.allclose(m.cell_volume[m.cell_zones["hex"]], h, rtol=1e-13)
    assert np.allclose(m.cell_volume[m.cell_zones["tet"]], h / 12, rtol=1e-12)
    assert abs(m.cell_volume.sum() - Lx * Ly * Lz) <= 1e-13 * Lx * Ly * Lz
    areas = {nm: m.face_area[f].sum() for nm, f in m.boundary_patches.items()}
    want = {"inlet": Ly * Lz, "outlet": Ly * Lz, "wall_bottom": Lx * Lz, "wall_top": Lx * Lz,
            "wall_front": Lx * Ly, "wall_back": Lx * Ly}
    assert set(areas) == set(want)
    for nm in want:
        assert abs(areas[nm] - want[nm]) <= 1e-13 * want[nm]
    # boundary faces on the outlet are triangles (two per hex face), on the inlet quads
    assert len(m.boundary_patches["outlet"]) == 2 * ny * nz
    assert len(m.boundary_patches["inlet"]) == ny * nz
    S = np.zeros((m.n_cells, 3))
    Sf = m.face_normal * m.face_area[:, None]
    np.add.at(S, m.owner, Sf)
    inner = m.neighbour >= 0
    np.add.at(S, m.neighbour[inner], -Sf[inner])
    assert np.max(np.abs(S)) <= 1e-15
    assert m.n_boundary_faces == sum(len(f) for f in m.boundary_patches.values())
