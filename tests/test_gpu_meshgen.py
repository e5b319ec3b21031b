"""The device mesh builder (kcfd_mesh_build: build_fvmesh_from_arrays and
_build_fvmesh_3d, mesh/mesh_reader.py:141-323, on the GPU).

Every array of the device-built FVMesh equals the host builder's bit for
bit: owner, neighbour, face nodes, boundary patches (names, order, faces),
face areas, normals and centres, cell centres and volumes.  The host
builder (k-cfd_amd/mesh.py) is itself pinned to the literal transcription
of the reference (tests/test_oracle_mesh.py).  Cases: 2D quads, hex (up to
128^3, the pressure-correction benchmark's mesh), Delaunay tets in random
cell order (the two-fluid benchmark's generator), the hybrid hex/tet duct
(ragged faces), boundary lists that overlap (the later row wins) or leave
faces unnamed ("default"), and the argument checks."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(md, mh, label):
    assert md.n_cells == mh.n_cells and md.n_faces == mh.n_faces, label
    assert md.n_internal_faces == mh.n_internal_faces and md.n_boundary_faces == mh.n_boundary_faces, label
    for f in ("owner", "neighbour", "face_nodes", "face_area", "face_normal", "face_center", "cell_center",
              "cell_volume"):
        a, b = getattr(md, f), getattr(mh, f)
        assert a.shape == b.shape, (label, f, a.shape, b.shape)
        if not np.array_equal(a, b):
            d = np.max(np.abs(np.asarray(a, float) - np.asarray(b, float)))
            raise AssertionError(f"{label}: {f} differs (max abs {d:.3e})")
    assert list(md.boundary_patches) == list(mh.boundary_patches), label
    for nm in mh.boundary_patches:
        assert np.array_equal(md.boundary_patches[nm], mh.boundary_patches[nm]), (label, nm)


def _both(label, *args, **kw):
    from kcfd_amd import mesh as M
    t = time.perf_counter()
    mh = M.build_fvmesh_from_arrays(*args, **kw)
    th = time.perf_counter() - t
    t = time.perf_counter()
    md = M.build_fvmesh_on_device(*args, **kw)
    td = time.perf_counter() - t
    print(f"{label}: {mh.n_cells} cells, {mh.n_faces} faces; host {th:.3f} s, device {td:.3f} s "
          f"({md.build_ms:.1f} ms on the device)")
    _same(md, mh, label)
    return md, mh


@pytest.mark.parametrize("n", [1, 7, 64])
def test_quad(ctx, n):
    from kcfd_amd import mesh as M
    nodes, cells, bf = M.structured_quad_arrays(0.0, 0.0, 1.0, 2.0, n, n + 3)
    _both(f"quad {n}x{n + 3}", nodes, cells, bf, ndim=2)


@pytest.mark.parametrize("n", [1, 5, 24])
def test_hex(ctx, n):
    from kcfd_amd import mesh as M
    nodes, cells, faces, bf = M.structured_hex_arrays(1.0, 0.5, 2.0, n, n + 1, n + 2)
    _both(f"hex {n}x{n + 1}x{n + 2}", nodes, cells, bf, ndim=3, cell_face_node_list=faces)


def test_hex128(ctx):
    from kcfd_amd import mesh as M
    nodes, cells, faces, bf = M.structured_hex_arrays(1.0, 1.0, 1.0, 128, 128, 128)
    md, _ = _both("hex 128^3", nodes, cells, bf, ndim=3, cell_face_node_list=faces)
    assert md.n_faces == 3 * 128 * 128 * 129


def test_tets_random_order(ctx):
    """The two-fluid benchmark's Delaunay generator (jittered lattice, slivers
    dropped, cells in random order) at 20,000 points: every face is met in
    an arbitrary cell order, and no boundary lists ("default")."""
    from scipy.spatial import Delaunay
    from kcfd_amd import mesh as M
    rng = np.random.default_rng(7)
    m = 27
    h = 1.0 / (m - 1)
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pts = idx * h + rng.uniform(-0.3, 0.3, (len(idx), 3)) * h * ((idx > 0) & (idx < m - 1))
    tets = Delaunay(pts).simplices.astype(np.int64)
    tets = tets[rng.permutation(len(tets))]
    md, mh = _both("tets", pts, tets, {}, ndim=3, cell_face_node_list=M.tet_arrays(tets))
    assert list(mh.boundary_patches) == ["default"]


def test_hybrid_ragged(ctx):
    from kcfd_amd import mesh as M
    nodes, rag, bdict, zones = M.hybrid_hex_tet_arrays(2.0, 0.1, 0.1, 24, 6, 6, 0.5)
    md, mh = _both("hybrid", nodes, None, bdict, cell_zone_map=zones, ndim=3, cell_face_node_list=rag)
    assert md.cell_zones == mh.cell_zones


def test_overlapping_and_missing_boundary_lists(ctx):
    """Patch names: a later list row re-names a face (here the first half
    of 'bottom' is listed again under 'lid'), faces in no list go to
    'default', an empty list names nothing, and patches are ordered by their
    first face."""
    from kcfd_amd import mesh as M
    nodes, cells, bf = M.structured_quad_arrays(0.0, 0.0, 1.0, 1.0, 9, 5)
    names = list(bf)
    b = {"empty": np.empty((0, 2), np.int64)}
    b[names[0]] = bf[names[0]]
    b[names[1]] = bf[names[1]]
    b["lid"] = np.asarray(bf[names[0]])[: len(bf[names[0]]) // 2][::-1, ::-1]   # reversed node order
    # names[2:] left out: those faces are 'default'
    md, mh = _both("quad overlapping lists", nodes, cells, b, ndim=2)
    assert "default" in mh.boundary_patches and "lid" in mh.boundary_patches
    assert "empty" not in mh.boundary_patches


def test_argument_checks(ctx):
    from kcfd_amd import mesh as M, _lib as L
    nodes, cells, bf = M.structured_quad_arrays(0.0, 0.0, 1.0, 1.0, 3, 3)
    bad = np.array(cells)
    bad[2, 1] = len(nodes)                 # node id out of range
    with pytest.raises(L.KcfdError, match="out of range"):
        M.build_fvmesh_on_device(nodes, bad, bf, ndim=2)
    hn, hc, hf, hb = M.structured_hex_arrays(1.0, 1.0, 1.0, 2, 2, 2)
    faces = np.array(hf)
    faces[0, 0, 2] = -1                    # a quad face with padding inside
    with pytest.raises(L.KcfdError, match="padding"):
        M.build_fvmesh_on_device(hn, hc, hb, ndim=3, cell_face_node_list=faces)


def _gen_both(label, host_args, dev_fn, dev_args, kind):
    from kcfd_amd import mesh as M
    t = time.perf_counter()
    if kind == "quad":
        nodes, cells, bf = M.structured_quad_arrays(*host_args)
        mh = M.build_fvmesh_from_arrays(nodes, cells, bf, ndim=2)
    else:
        nodes, cells, faces, bf = M.structured_hex_arrays(*host_args)
        mh = M.build_fvmesh_from_arrays(nodes, cells, bf, ndim=3, cell_face_node_list=faces)
    th = time.perf_counter() - t
    t = time.perf_counter()
    md = dev_fn(*dev_args)
    td = time.perf_counter() - t
    print(f"{label}: {mh.n_cells} cells; host generate + build {th:.3f} s, device {td:.3f} s "
          f"({md.build_ms:.1f} ms on the device)")
    assert md.nodes.shape == mh.nodes.shape and np.array_equal(md.nodes, mh.nodes), label
    _same(md, mh, label)
    return md, mh


@pytest.mark.parametrize("n", [1, 9, 128])
def test_generated_quad(ctx, n):
    from kcfd_amd import mesh as M
    args = (0.25, -1.0, 1.75, 2.0, n, n + 5)
    _gen_both(f"generated quad {n}x{n + 5}", args, M.structured_quad_mesh_on_device, args, "quad")


@pytest.mark.parametrize("n", [1, 6, 128])
def test_generated_hex(ctx, n):
    from kcfd_amd import mesh as M
    args = (1.0, 0.5, 2.0, n, max(1, n // 2), n + 1) if n < 128 else (1.0, 1.0, 1.0, n, n, n)
    _gen_both(f"generated hex {args[3:]}", args, M.structured_hex_mesh_on_device, args, "hex")


def test_generated_shared_patch_names(ctx):
    """Sides given the same name: the generators' dict keeps the name at its
    first side's position with the last side's faces; the earlier sides'
    faces then fall to 'default'."""
    from kcfd_amd import mesh as M
    names = {"x_min": "in", "x_max": "out", "y_min": "walls", "y_max": "walls", "z_min": "walls",
             "z_max": "walls"}
    args = (1.0, 1.0, 1.0, 5, 4, 3)
    md, mh = _gen_both("hex shared names", args + (names,), M.structured_hex_mesh_on_device, args + (names,), "hex")
    assert "default" in mh.boundary_patches
    qn = {"bottom": "wall", "top": "wall", "left": "in"}
    qargs = (0.0, 0.0, 1.0, 1.0, 7, 4, qn)
    _gen_both("quad shared names", qargs, M.structured_quad_mesh_on_device, qargs, "quad")
