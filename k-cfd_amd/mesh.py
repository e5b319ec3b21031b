"""Structure-of-arrays finite-volume mesh and vectorised builders.

The reference keeps its mesh as Python objects per face and per cell
(mesh/mesh_reader.py:13-63) and builds them with dictionary loops
(mesh/mesh_reader.py:141-323).  That layout cannot reach the 2M-cell hex or
1.1M-cell tet workloads, so this module rebuilds the same mesh with numpy in
structure-of-arrays form.  The rules that the assembly depends on are kept
exactly:

* faces are numbered in first-encounter order while walking cells in order and
  each cell's edges/faces in order (mesh_reader.py:189-214, :269-290);
* the first cell that meets a face is its owner, the second its neighbour;
* face normals point out of the owner (mesh_reader.py:204-206, :280-282);
* 2D face: midpoint centre, normal (dy,-dx)/L, area L (mesh_reader.py:66-78);
* 3D face: vertex-mean centre, Newell normal, area |N|/2 (mesh_reader.py:81-99);
* 2D cell: vertex-mean centre, shoelace area (mesh_reader.py:102-113);
* 3D cell: vertex-mean centre, |sum sign*(x_f.n_f)A_f|/3 floored at 1e-30
  (mesh_reader.py:116-138);
* boundary patches list face ids in face order; a boundary face whose key is
  not in the patch dictionary goes to "default"; patches appear in the order
  their first face is met (mesh_reader.py:224-231, :306-313).

Geometry is stored padded to three components (z = 0 in 2D), which leaves
every norm and dot product of the reference unchanged.
"""

from __future__ import annotations

import numpy as np
from typing import NamedTuple

__all__ = ["FVMesh", "build_fvmesh_from_arrays", "structured_quad_arrays",
           "structured_hex_arrays", "tet_arrays", "RaggedCells", "hybrid_hex_tet_arrays",
           "generate_hybrid_hex_tet_mesh", "build_fvmesh_on_device", "structured_quad_mesh_on_device",
           "structured_hex_mesh_on_device"]


class FVMesh:
    """Cell-centred FV mesh in SoA form (same public counters as the reference
    ``FVMesh``, mesh/mesh_reader.py:34-52)."""

    def __init__(self, ndim: int):
        self.ndim = ndim
        self.nodes = np.empty((0, ndim))
        self.n_cells = 0
        self.n_faces = 0
        self.n_internal_faces = 0
        self.n_boundary_faces = 0
        self.boundary_patches: dict[str, np.ndarray] = {}
        self.cell_zones: dict = {}
        # per face
        self.owner = np.empty(0, np.int32)
        self.neighbour = np.empty(0, np.int32)
        self.face_area = np.empty(0)
        self.face_normal = np.empty((0, 3))
        self.face_center = np.empty((0, 3))
        # per cell
        self.cell_center = np.empty((0, 3))
        self.cell_volume = np.empty(0)
        self._derived = None

    # ---- derived, cached quantities -------------------------------------
    def derived(self) -> dict:
        """Face/cell quantities the assembly, gradient and MUSCL kernels read.

        All are computed with the reference's own formulas:
        d_PN = |x_N - x_O| (fvm_operators.py:77), d_Pf = |x_f - x_O|
        (fvm_operators.py:197), g = d_N/(d_O+d_N) or 0.5
        (gradient.py:102-114)."""
        if self._derived is not None:
            return self._derived
        o = self.owner.astype(np.int64)
        nb = self.neighbour.astype(np.int64)
        internal = nb >= 0
        xo = self.cell_center[o]
        xn = self.cell_center[np.where(internal, nb, o)]
        xf = self.face_center
        d_pn = np.sqrt(np.sum((xn - xo) ** 2, axis=1))
        d_pf = np.sqrt(np.sum((xf - xo) ** 2, axis=1))
        d_nf = np.sqrt(np.sum((xf - xn) ** 2, axis=1))
        tot = d_pf + d_nf
        with np.errstate(invalid="ignore", divide="ignore"):
            g = np.where(tot < 1e-30, 0.5, d_nf / np.where(tot < 1e-30, 1.0, tot))
        g = np.where(internal, g, 1.0)
        # boundary face -> (patch index, local index)
        bpatch = np.full(self.n_faces, -1, np.int32)
        blocal = np.full(self.n_faces, -1, np.int32)
        for pi, (name, fids) in enumerate(self.boundary_patches.items()):
            bpatch[fids] = pi
            blocal[fids] = np.arange(len(fids), dtype=np.int32)
        # cell -> faces CSR in face-id order, with the side (+1 owner, -1 neighbour)
        cf_cell = np.concatenate([o, nb[internal]])
        cf_face = np.concatenate([np.arange(self.n_faces), np.nonzero(internal)[0]])
        cf_side = np.concatenate([np.ones(self.n_faces, np.int8),
                                  -np.ones(int(internal.sum()), np.int8)])
        order = np.lexsort((cf_face, cf_cell))
        cf_cell, cf_face, cf_side = cf_cell[order], cf_face[order], cf_side[order]
        cf_ptr = np.zeros(self.n_cells + 1, np.int64)
        np.add.at(cf_ptr, cf_cell + 1, 1)
        cf_ptr = np.cumsum(cf_ptr)
        self._derived = dict(d_pn=d_pn, d_pf=d_pf, d_nf=d_nf, gweight=g,
                             internal=internal, bpatch=bpatch, blocal=blocal,
                             cf_ptr=cf_ptr.astype(np.int32),
                             cf_face=cf_face.astype(np.int32),
                             cf_side=cf_side)
        return self._derived

    @property
    def patch_names(self) -> list:
        return list(self.boundary_patches.keys())

    def summary(self) -> str:
        s = (f"FVMesh({self.ndim}D): {self.n_cells} cells, {self.n_faces} faces "
             f"({self.n_internal_faces} internal, {self.n_boundary_faces} boundary)\n")
        for name, fids in self.boundary_patches.items():
            s += f"  Boundary '{name}': {len(fids)} faces\n"
        return s


# node id standing for "no node" in sorted face keys: sorts after every real
# id and keeps the keys non-negative (the stored face arrays pad with -1)
_NO_NODE = (1 << 31) - 1


class RaggedCells(NamedTuple):
    """3D cells with differing node and face counts (mixed hex/tet meshes).

    cell_nodes: (n, 8) node ids, padded with -1; faces: (E, 4) the faces of
    every cell in cell order, each padded with -1 (triangles); face_ptr:
    (n + 1) offsets of each cell's faces in ``faces``."""
    cell_nodes: np.ndarray
    faces: np.ndarray
    face_ptr: np.ndarray


def _face_keys(rows: np.ndarray) -> np.ndarray:
    """Sorted node keys of padded face rows (-1 -> _NO_NODE, so a triangle
    never equals a quad)."""
    k = np.where(rows >= 0, rows, _NO_NODE).astype(np.int64)
    k.sort(axis=1)
    return k


def _face_rows(flist, width: int) -> np.ndarray:
    """A boundary face list (array, or list of node arrays of any length up
    to width) as padded (m, width) rows."""
    if isinstance(flist, np.ndarray) and flist.ndim == 2:
        if flist.shape[1] == width:
            return flist.astype(np.int64)
        out = np.full((len(flist), width), -1, np.int64)
        out[:, :flist.shape[1]] = flist
        return out
    out = np.full((len(flist), width), -1, np.int64)
    for r, f in enumerate(flist):
        f = np.asarray(f, dtype=np.int64).ravel()
        out[r, :len(f)] = f
    return out


def _pad3(a: np.ndarray) -> np.ndarray:
    if a.shape[1] == 3:
        return np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros((a.shape[0], 3))
    out[:, : a.shape[1]] = a
    return out


def _pack_keys(keys: np.ndarray) -> list:
    """Pack rows of small non-negative ints into at most two int64 words
    (32 bits per node id), most significant first."""
    k = keys.astype(np.int64)
    w = k.shape[1]
    if w <= 2:
        return [(k[:, 0] << 32) | k[:, 1]] if w == 2 else [k[:, 0]]
    if w <= 4:
        hi = (k[:, 0] << 32) | k[:, 1]
        lo = (k[:, 2] << 32) | (k[:, 3] if w == 4 else 0)
        return [hi, lo]
    raise ValueError("faces with more than 4 nodes are not supported")


def _first_encounter_faces(keys: np.ndarray):
    """Number unique rows of ``keys`` by first occurrence.

    Returns (face_of_entry, first_entry_of_face, last_other_entry_of_face or -1)."""
    nk = keys.shape[0]
    words = _pack_keys(keys)
    order = np.lexsort(words[::-1])          # stable: ties keep entry order
    same = np.zeros(nk, bool)          # sorted entry equals the previous one
    if nk > 1:
        same[1:] = True
        for wv in words:
            sv = wv[order]
            same[1:] &= sv[1:] == sv[:-1]
    start = ~same
    gid_sorted = np.cumsum(start) - 1
    ng = int(gid_sorted[-1]) + 1 if nk else 0
    starts = np.nonzero(start)[0]
    ends = np.append(starts[1:], nk) - 1
    first_of_group = order[starts]
    last_of_group = order[ends]
    # faces are numbered by first occurrence
    rank_order = np.argsort(first_of_group, kind="stable")
    rank = np.empty(ng, np.int64)
    rank[rank_order] = np.arange(ng)
    face_of_entry = np.empty(nk, np.int64)
    face_of_entry[order] = rank[gid_sorted]
    first_entry = first_of_group[rank_order]
    last_entry = last_of_group[rank_order]
    # the last entry other than the first that meets the face is its neighbour
    # (later encounters overwrite .neighbour, mesh_reader.py:198,275)
    second = np.where(last_entry != first_entry, last_entry, -1)
    return face_of_entry, first_entry, second


def _tag_boundaries(mesh: FVMesh, face_keys: np.ndarray, boundary_faces_dict: dict,
                    width: int):
    bmask = mesh.neighbour < 0
    bfids = np.nonzero(bmask)[0]
    names = list(boundary_faces_dict.keys())
    if len(names):
        bk = []
        bname_idx = []
        for ni, (name, flist) in enumerate(boundary_faces_dict.items()):
            if len(flist) == 0:
                continue
            arr = _face_keys(_face_rows(flist, width))
            bk.append(arr)
            bname_idx.append(np.full(len(arr), ni, np.int64))
        bk = np.concatenate(bk) if bk else np.empty((0, width), np.int64)
        bname_idx = np.concatenate(bname_idx) if bname_idx else np.empty(0, np.int64)
    else:
        bk = np.empty((0, width), np.int64)
        bname_idx = np.empty(0, np.int64)
    # later dictionary entries overwrite earlier ones for the same key
    # (boundary_edge_set[ek] = bname, mesh_reader.py:218-222)
    fk = face_keys[bfids]
    allk = np.concatenate([bk, fk])
    kview = np.ascontiguousarray(allk).view(
        np.dtype((np.void, allk.dtype.itemsize * width))).ravel()
    _, inv = np.unique(kview, return_inverse=True)
    inv = inv.ravel()
    nb = len(bk)
    key_name = np.full(inv.max() + 1 if len(inv) else 0, -1, np.int64)
    key_name[inv[:nb]] = bname_idx  # fancy assignment: last write wins
    face_name = key_name[inv[nb:]]
    all_names = names + ["default"]
    face_name = np.where(face_name < 0, len(names), face_name)
    # patch order: order of first boundary face per name
    patches = {}
    if len(face_name):
        uniq_names, first_pos = np.unique(face_name, return_index=True)
        for nm_idx in uniq_names[np.argsort(first_pos)]:
            patches[all_names[nm_idx]] = bfids[face_name == nm_idx].astype(np.int64)
    mesh.boundary_patches = patches


def build_fvmesh_from_arrays(node_coords, cell_node_list, boundary_faces_dict,
                             cell_zone_map=None, ndim: int = 2,
                             cell_face_node_list=None) -> FVMesh:
    """Vectorised equivalent of ``build_fvmesh_from_arrays``
    (mesh/mesh_reader.py:141-242) and ``_build_fvmesh_3d`` (:245-323).

    ``cell_node_list`` may be a 2-D int array (uniform cell type) or a list.
    In 3D, ``cell_face_node_list`` may be an int array of shape
    (n_cells, faces_per_cell, nodes_per_face), a ``RaggedCells`` (mixed cell
    types; ``cell_node_list`` is then ignored), or the reference's list of
    per-cell face lists, uniform or not."""
    nodes = np.asarray(node_coords, dtype=np.float64)
    if ndim == 3 and cell_face_node_list is not None:
        rag = _as_ragged(cell_node_list, cell_face_node_list)
        if rag is not None:
            return _build_3d_ragged(nodes, rag, boundary_faces_dict, cell_zone_map)
    mesh = FVMesh(ndim)
    mesh.nodes = nodes
    cells = np.asarray(cell_node_list, dtype=np.int64)
    n = cells.shape[0]
    mesh.n_cells = n
    nodes3 = _pad3(nodes)
    if ndim == 3 and cell_face_node_list is not None:
        cfn = np.asarray(cell_face_node_list, dtype=np.int64)
        _, fpc, npf = cfn.shape
        entries = cfn.reshape(n * fpc, npf)
        keys = np.sort(entries, axis=1)
        face_of_entry, first, second = _first_encounter_faces(keys)
        ent_cell = np.repeat(np.arange(n), fpc)
        fnodes = entries[first]
        owner = ent_cell[first]
        neighbour = np.where(second >= 0, ent_cell[np.maximum(second, 0)], -1)
        pts = nodes3[fnodes]                       # (nf, npf, 3)
        center = pts.mean(axis=1)
        nxt = np.roll(pts, -1, axis=1)
        nrm = np.zeros((len(first), 3))
        nrm[:, 0] = np.sum((pts[:, :, 1] - nxt[:, :, 1]) * (pts[:, :, 2] + nxt[:, :, 2]), axis=1)
        nrm[:, 1] = np.sum((pts[:, :, 2] - nxt[:, :, 2]) * (pts[:, :, 0] + nxt[:, :, 0]), axis=1)
        nrm[:, 2] = np.sum((pts[:, :, 0] - nxt[:, :, 0]) * (pts[:, :, 1] + nxt[:, :, 1]), axis=1)
        area = np.sqrt(np.sum(nrm * nrm, axis=1))
        nz = area > 1e-15
        nrm[nz] /= area[nz, None]
        area = area * 0.5
        ccen = nodes3[cells].mean(axis=1)
        to_face = center - ccen[owner]
        flip = np.sum(nrm * to_face, axis=1) < 0
        nrm[flip] = -nrm[flip]
        # cell volume by the divergence theorem over the cell's own face list
        ef = face_of_entry
        ec = ent_cell
        tf = center[ef] - ccen[ec]
        sgn = np.where(np.sum(nrm[ef] * tf, axis=1) < 0, -1.0, 1.0)
        contrib = sgn * np.sum(center[ef] * nrm[ef], axis=1) * area[ef]
        vol = np.zeros(n)
        # sequential per-cell sum in the cell's face order
        vol = np.add.reduceat(contrib, np.arange(0, n * fpc, fpc)) if n else vol
        vol = np.abs(vol) / 3.0
        vol = np.where(vol < 1e-30, 1e-30, vol)
        width = npf
    else:
        npc = cells.shape[1]
        e0 = cells
        e1 = np.roll(cells, -1, axis=1)
        entries = np.stack([e0, e1], axis=2).reshape(n * npc, 2)
        keys = np.sort(entries, axis=1)
        face_of_entry, first, second = _first_encounter_faces(keys)
        ent_cell = np.repeat(np.arange(n), npc)
        fnodes = entries[first]
        owner = ent_cell[first]
        neighbour = np.where(second >= 0, ent_cell[np.maximum(second, 0)], -1)
        p0 = nodes3[fnodes[:, 0]]
        p1 = nodes3[fnodes[:, 1]]
        center = 0.5 * (p0 + p1)
        dx = p1[:, 0] - p0[:, 0]
        dy = p1[:, 1] - p0[:, 1]
        length = np.sqrt(dx ** 2 + dy ** 2)
        nrm = np.zeros((len(first), 3))
        nrm[:, 0] = dy
        nrm[:, 1] = -dx
        nz = length > 1e-15
        nrm[nz, :2] /= length[nz, None]
        area = length
        pts = nodes3[cells]
        ccen = pts.mean(axis=1)
        nx_ = np.roll(pts, -1, axis=1)
        sh = np.zeros(n)
        for k in range(npc):  # sequential shoelace sum in node order
            sh = sh + (pts[:, k, 0] * nx_[:, k, 1] - nx_[:, k, 0] * pts[:, k, 1])
        vol = np.abs(sh) / 2.0
        to_face = center - ccen[owner]
        flip = np.sum(nrm * to_face, axis=1) < 0
        nrm[flip] = -nrm[flip]
        width = 2
    mesh.owner = owner.astype(np.int32)
    mesh.neighbour = neighbour.astype(np.int32)
    mesh.face_area = np.ascontiguousarray(area)
    mesh.face_normal = np.ascontiguousarray(nrm)
    mesh.face_center = np.ascontiguousarray(center)
    mesh.cell_center = np.ascontiguousarray(ccen)
    mesh.cell_volume = np.ascontiguousarray(vol)
    mesh.face_nodes = fnodes
    mesh.n_faces = len(first)
    mesh.n_internal_faces = int(np.sum(neighbour >= 0))
    mesh.n_boundary_faces = mesh.n_faces - mesh.n_internal_faces
    _tag_boundaries(mesh, np.sort(fnodes, axis=1), boundary_faces_dict or {}, width)
    if cell_zone_map:
        mesh.cell_zones = cell_zone_map
    return mesh


def build_fvmesh_on_device(node_coords, cell_node_list, boundary_faces_dict, cell_zone_map=None,
                           ndim: int = 2, cell_face_node_list=None, ctx=None) -> FVMesh:
    """``build_fvmesh_from_arrays`` (mesh/mesh_reader.py:141-323) with the
    face matching and all geometry computed on the GPU (kcfd_mesh_build):
    the same arguments, and the same FVMesh as the host builder above, array
    for array and bit for bit.  ``ctx``: the library context (default: the
    process's).  The build's device time is kept as ``mesh.build_ms``."""
    from . import _lib as L
    ctx = ctx or L.default_context()
    nodes = np.asarray(node_coords, dtype=np.float64)
    nodes3 = _pad3(nodes.reshape(len(nodes), -1))
    mode = 2
    entries = ptr = None
    if ndim == 3 and cell_face_node_list is not None:
        mode = 3
        rag = _as_ragged(cell_node_list, cell_face_node_list)
        if rag is not None:
            cells = np.asarray(rag.cell_nodes, dtype=np.int64)
            entries = np.asarray(rag.faces, dtype=np.int64)
            ptr = np.asarray(rag.face_ptr, dtype=np.int64)
            width = 4
        else:
            cells = np.asarray(cell_node_list, dtype=np.int64)
            cfn = np.asarray(cell_face_node_list, dtype=np.int64)
            nc, fpc, width = cfn.shape
            entries = cfn.reshape(nc * fpc, width)
            ptr = np.arange(0, nc * fpc + 1, fpc, dtype=np.int64)
    else:
        cells = np.asarray(cell_node_list, dtype=np.int64)
        width = 2
    n = cells.shape[0]
    names = list((boundary_faces_dict or {}).keys())
    brows, bname = [], []
    for ni, (name, flist) in enumerate((boundary_faces_dict or {}).items()):
        if len(flist) == 0:
            continue
        rows = _face_rows(flist, width)
        brows.append(rows)
        bname.append(np.full(len(rows), ni, np.int64))
    brows = np.concatenate(brows) if brows else np.empty((0, width), np.int64)
    bname = np.concatenate(bname) if bname else np.empty(0, np.int64)

    def i32(a):
        return np.ascontiguousarray(a, dtype=np.int32)

    cells32 = i32(cells)
    ent32 = i32(entries) if entries is not None else None
    ptr32 = i32(ptr) if ptr is not None else None
    b32, bn32 = i32(brows), i32(bname)
    h = L.ctypes.c_void_p()
    L.check(L.lib().kcfd_mesh_build(
        ctx.h, mode, len(nodes3), L.ptr(nodes3), n, cells.shape[1] if n else 1, L.ptr(cells32),
        len(ent32) if ent32 is not None else 0, width, L.ptr(ent32) if ent32 is not None else None,
        L.ptr(ptr32) if ptr32 is not None else None, len(b32), L.ptr(b32) if len(b32) else None,
        L.ptr(bn32) if len(bn32) else None, len(names), L.ctypes.byref(h)))
    return _mesh_of_faces(h, ndim, nodes, names, cell_zone_map)


def _mesh_of_faces(h, ndim, nodes, names, cell_zone_map=None) -> FVMesh:
    """The FVMesh of a device-built face set (frees the handle).  ``nodes``:
    the node array to keep (None: the device's, with ndim columns)."""
    from . import _lib as L
    try:
        nf, nint, fw = L.ctypes.c_int(), L.ctypes.c_int(), L.ctypes.c_int()
        nn, nc = L.ctypes.c_int(), L.ctypes.c_int()
        ms = L.ctypes.c_double()
        L.check(L.lib().kcfd_mesh_faces_info(h, L.ctypes.byref(nf), L.ctypes.byref(nint), L.ctypes.byref(fw),
                                             L.ctypes.byref(ms)))
        L.check(L.lib().kcfd_mesh_faces_nodes(h, L.ctypes.byref(nn), L.ctypes.byref(nc), None))
        if nodes is None:
            xyz = np.empty((nn.value, 3))
            L.check(L.lib().kcfd_mesh_faces_nodes(h, None, None, L.ptr(xyz)))
            nodes = np.ascontiguousarray(xyz[:, :ndim])
        nf, n = nf.value, nc.value
        owner = np.empty(nf, np.int32)
        neighbour = np.empty(nf, np.int32)
        fnodes = np.empty((nf, fw.value), np.int32)
        fpatch = np.empty(nf, np.int32)
        area = np.empty(nf)
        nrm = np.empty((nf, 3))
        fc = np.empty((nf, 3))
        cc = np.empty((n, 3))
        vol = np.empty(n)
        L.check(L.lib().kcfd_mesh_faces_get(h, L.ptr(owner), L.ptr(neighbour), L.ptr(fnodes), L.ptr(fpatch),
                                            L.ptr(area), L.ptr(nrm), L.ptr(fc), L.ptr(cc), L.ptr(vol)))
    finally:
        L.lib().kcfd_mesh_faces_free(h)
    mesh = FVMesh(ndim)
    mesh.nodes = nodes
    mesh.n_cells = n
    mesh.owner, mesh.neighbour = owner, neighbour
    mesh.face_area, mesh.face_normal, mesh.face_center = area, nrm, fc
    mesh.cell_center, mesh.cell_volume = cc, vol
    mesh.face_nodes = fnodes.astype(np.int64)
    mesh.n_faces = nf
    mesh.n_internal_faces = int(nint.value)
    mesh.n_boundary_faces = nf - mesh.n_internal_faces
    # patches in the order of their first boundary face ("default" last index)
    bfids = np.nonzero(neighbour < 0)[0]
    fp = fpatch[bfids]
    all_names = list(names) + ["default"]
    patches = {}
    if len(fp):
        uniq, first = np.unique(fp, return_index=True)
        for idx in uniq[np.argsort(first)]:
            patches[all_names[idx]] = bfids[fp == idx].astype(np.int64)
    mesh.boundary_patches = patches
    if cell_zone_map:
        mesh.cell_zones = cell_zone_map
    mesh.build_ms = ms.value
    return mesh


def _side_patches(keys, defaults, boundary_names):
    """Patch names in the generators' dict order and, per side, the index of
    the patch it is listed under (-1 where a later side of the same name
    replaces its list, as the generators' dict assignment does)."""
    names, side_name = [], []
    for key, dflt in zip(keys, defaults):
        nm = boundary_names.get(key, dflt)
        side_name.append(nm)
        if nm not in names:
            names.append(nm)
    last = {nm: q for q, nm in enumerate(side_name)}
    side = np.array([names.index(nm) if last[nm] == q else -1 for q, nm in enumerate(side_name)], np.int32)
    return names, side


def structured_quad_mesh_on_device(x0, y0, x1, y1, nx, ny, boundary_names=None, ctx=None) -> FVMesh:
    """The mesh of ``structured_quad_arrays`` + ``build_fvmesh_from_arrays``
    (mesh/mesh_generator.py:13-93, mesh/mesh_reader.py:141-242), generated
    and built on the GPU (kcfd_mesh_generate): bit for bit the host's."""
    from . import _lib as L
    ctx = ctx or L.default_context()
    keys = ["bottom", "top", "left", "right"]
    names, side = _side_patches(keys, keys, boundary_names or {k: k for k in keys})
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    h = L.ctypes.c_void_p()
    L.check(L.lib().kcfd_mesh_generate(ctx.h, 2, int(nx), int(ny), 0, float(x0), float(y0), float(dx), float(dy),
                                       0.0, L.ptr(side), len(names), L.ctypes.byref(h)))
    return _mesh_of_faces(h, 2, None, names)


def structured_hex_mesh_on_device(Lx, Ly, Lz, nx, ny, nz, boundary_names=None, ctx=None) -> FVMesh:
    """The mesh of ``structured_hex_arrays`` + ``build_fvmesh_from_arrays``
    (mesh/mesh_generator_3d.py:12-159, mesh/mesh_reader.py:245-323),
    generated and built on the GPU (kcfd_mesh_generate): bit for bit the
    host's."""
    from . import _lib as L
    ctx = ctx or L.default_context()
    keys = ["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"]
    dflt = ["inlet", "outlet", "wall_bottom", "wall_top", "wall_front", "wall_back"]
    if boundary_names is None:
        boundary_names = dict(zip(keys, dflt))
    names, side = _side_patches(keys, dflt, boundary_names)
    h = L.ctypes.c_void_p()
    L.check(L.lib().kcfd_mesh_generate(ctx.h, 3, int(nx), int(ny), int(nz), 0.0, 0.0, float(Lx / nx),
                                       float(Ly / ny), float(Lz / nz), L.ptr(side), len(names),
                                       L.ctypes.byref(h)))
    return _mesh_of_faces(h, 3, None, names)


def _as_ragged(cell_node_list, cell_face_node_list):
    """RaggedCells for mixed 3D cells, or None when the per-cell face lists
    form one uniform (n, faces, nodes) array."""
    if isinstance(cell_face_node_list, RaggedCells):
        return cell_face_node_list
    if isinstance(cell_face_node_list, np.ndarray):
        return None
    try:
        np.asarray(cell_face_node_list, dtype=np.int64)
        np.asarray(cell_node_list, dtype=np.int64)
        return None
    except ValueError:
        pass
    n = len(cell_face_node_list)
    cn = np.full((n, 8), -1, np.int64)
    for c, nodes in enumerate(cell_node_list):
        v = np.asarray(nodes, dtype=np.int64).ravel()
        cn[c, :len(v)] = v
    counts = np.array([len(fl) for fl in cell_face_node_list], np.int64)
    ptr = np.concatenate([[0], np.cumsum(counts)])
    faces = _face_rows([f for fl in cell_face_node_list for f in fl], 4)
    return RaggedCells(cn, faces, ptr)


def _build_3d_ragged(nodes, rag: RaggedCells, boundary_faces_dict, cell_zone_map) -> FVMesh:
    """``_build_fvmesh_3d`` (mesh/mesh_reader.py:245-323) for cells of
    differing node and face counts: the uniform path's arithmetic applied to
    each group of faces (3 or 4 nodes) and cells (4 or 8 nodes)."""
    cn = np.asarray(rag.cell_nodes, dtype=np.int64)
    fe = np.asarray(rag.faces, dtype=np.int64)
    fptr = np.asarray(rag.face_ptr, dtype=np.int64)
    n = cn.shape[0]
    mesh = FVMesh(3)
    mesh.nodes = nodes
    mesh.n_cells = n
    nodes3 = _pad3(nodes)
    keys = _face_keys(fe)
    face_of_entry, first, second = _first_encounter_faces(keys)
    ent_cell = np.repeat(np.arange(n), np.diff(fptr))
    fnodes = fe[first]
    owner = ent_cell[first]
    neighbour = np.where(second >= 0, ent_cell[np.maximum(second, 0)], -1)
    nf = len(first)
    fk = (fnodes >= 0).sum(axis=1)
    center = np.zeros((nf, 3))
    nrm = np.zeros((nf, 3))
    for k in np.unique(fk):
        sel = np.nonzero(fk == k)[0]
        pts = nodes3[fnodes[sel, :k]]
        center[sel] = pts.mean(axis=1)
        nxt = np.roll(pts, -1, axis=1)
        nrm[sel, 0] = np.sum((pts[:, :, 1] - nxt[:, :, 1]) * (pts[:, :, 2] + nxt[:, :, 2]), axis=1)
        nrm[sel, 1] = np.sum((pts[:, :, 2] - nxt[:, :, 2]) * (pts[:, :, 0] + nxt[:, :, 0]), axis=1)
        nrm[sel, 2] = np.sum((pts[:, :, 0] - nxt[:, :, 0]) * (pts[:, :, 1] + nxt[:, :, 1]), axis=1)
    area = np.sqrt(np.sum(nrm * nrm, axis=1))
    nz = area > 1e-15
    nrm[nz] /= area[nz, None]
    area = area * 0.5
    ccen = np.zeros((n, 3))
    cnn = (cn >= 0).sum(axis=1)
    for k in np.unique(cnn):
        sel = np.nonzero(cnn == k)[0]
        ccen[sel] = nodes3[cn[sel, :k]].mean(axis=1)
    to_face = center - ccen[owner]
    flip = np.sum(nrm * to_face, axis=1) < 0
    nrm[flip] = -nrm[flip]
    ef, ec = face_of_entry, ent_cell
    tf = center[ef] - ccen[ec]
    sgn = np.where(np.sum(nrm[ef] * tf, axis=1) < 0, -1.0, 1.0)
    contrib = sgn * np.sum(center[ef] * nrm[ef], axis=1) * area[ef]
    vol = np.add.reduceat(contrib, fptr[:-1]) if n else np.zeros(0)
    vol = np.abs(vol) / 3.0
    vol = np.where(vol < 1e-30, 1e-30, vol)
    mesh.owner = owner.astype(np.int32)
    mesh.neighbour = neighbour.astype(np.int32)
    mesh.face_area = np.ascontiguousarray(area)
    mesh.face_normal = np.ascontiguousarray(nrm)
    mesh.face_center = np.ascontiguousarray(center)
    mesh.cell_center = np.ascontiguousarray(ccen)
    mesh.cell_volume = np.ascontiguousarray(vol)
    mesh.face_nodes = fnodes
    mesh.n_faces = nf
    mesh.n_internal_faces = int(np.sum(neighbour >= 0))
    mesh.n_boundary_faces = nf - mesh.n_internal_faces
    _tag_boundaries(mesh, _face_keys(fnodes), boundary_faces_dict or {}, 4)
    if cell_zone_map:
        mesh.cell_zones = cell_zone_map
    return mesh


# ---------------------------------------------------------------------------
# generator inputs (node arrays in the reference generators' numbering)
# ---------------------------------------------------------------------------

def structured_quad_arrays(x0, y0, x1, y1, nx, ny, boundary_names=None):
    """Arrays of ``_make_structured_quad_mesh`` (mesh/mesh_generator.py:13-93):
    row-major nodes, counter-clockwise quads, patches bottom, top, left, right."""
    if boundary_names is None:
        boundary_names = {"bottom": "bottom", "top": "top", "left": "left", "right": "right"}
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    nodes = np.stack([x0 + ii.ravel() * dx, y0 + jj.ravel() * dy], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (j * (nx + 1) + i).ravel()
    cells = np.stack([n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)], axis=1)
    ar = np.arange(nx)
    br = np.arange(ny)
    bfaces = {
        boundary_names.get("bottom", "bottom"): np.stack([ar, ar + 1], 1),
        boundary_names.get("top", "top"): np.stack([ny * (nx + 1) + ar, ny * (nx + 1) + ar + 1], 1),
        boundary_names.get("left", "left"): np.stack([br * (nx + 1), (br + 1) * (nx + 1)], 1),
        boundary_names.get("right", "right"): np.stack([br * (nx + 1) + nx, (br + 1) * (nx + 1) + nx], 1),
    }
    return nodes, cells, bfaces


def structured_hex_arrays(Lx, Ly, Lz, nx, ny, nz, boundary_names=None):
    """Arrays of ``generate_3d_channel_mesh`` (mesh/mesh_generator_3d.py:12-159):
    node id k*(ny+1)*(nx+1)+j*(nx+1)+i, cells i fastest, six faces per cell in
    the order z-, z+, y-, y+, x-, x+."""
    if boundary_names is None:
        boundary_names = {"x_min": "inlet", "x_max": "outlet", "y_min": "wall_bottom",
                          "y_max": "wall_top", "z_min": "wall_front", "z_max": "wall_back"}
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    nodes = np.stack([ii.ravel() * dx, jj.ravel() * dy, kk.ravel() * dz], axis=1)

    def nid(i, j, k):
        return k * (ny + 1) * (nx + 1) + j * (nx + 1) + i

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    n0, n1, n2, n3 = nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k)
    n4, n5, n6, n7 = nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)
    cells = np.stack([n0, n1, n2, n3, n4, n5, n6, n7], axis=1)
    faces = np.stack([
        np.stack([n0, n3, n2, n1], 1), np.stack([n4, n5, n6, n7], 1),
        np.stack([n0, n1, n5, n4], 1), np.stack([n2, n3, n7, n6], 1),
        np.stack([n0, n4, n7, n3], 1), np.stack([n1, n2, n6, n5], 1)], axis=1)
    kb, jb = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    kb, jb = kb.ravel(), jb.ravel()
    kc, ic = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    kc, ic = kc.ravel(), ic.ravel()
    jd, idd = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    jd, idd = jd.ravel(), idd.ravel()
    bf = {}
    bf[boundary_names.get("x_min", "inlet")] = np.stack(
        [nid(0, jb, kb), nid(0, jb, kb + 1), nid(0, jb + 1, kb + 1), nid(0, jb + 1, kb)], 1)
    bf[boundary_names.get("x_max", "outlet")] = np.stack(
        [nid(nx, jb, kb), nid(nx, jb + 1, kb), nid(nx, jb + 1, kb + 1), nid(nx, jb, kb + 1)], 1)
    bf[boundary_names.get("y_min", "wall_bottom")] = np.stack(
        [nid(ic, 0, kc), nid(ic + 1, 0, kc), nid(ic + 1, 0, kc + 1), nid(ic, 0, kc + 1)], 1)
    bf[boundary_names.get("y_max", "wall_top")] = np.stack(
        [nid(ic + 1, ny, kc), nid(ic, ny, kc), nid(ic, ny, kc + 1), nid(ic + 1, ny, kc + 1)], 1)
    bf[boundary_names.get("z_min", "wall_front")] = np.stack(
        [nid(idd, jd, 0), nid(idd, jd + 1, 0), nid(idd + 1, jd + 1, 0), nid(idd + 1, jd, 0)], 1)
    bf[boundary_names.get("z_max", "wall_back")] = np.stack(
        [nid(idd, jd, nz), nid(idd + 1, jd, nz), nid(idd + 1, jd + 1, nz), nid(idd, jd + 1, nz)], 1)
    return nodes, cells, faces, bf


def tet_arrays(tets: np.ndarray):
    """Per-cell face node lists of a tetrahedral mesh in the order of
    ``_tet_faces`` (mesh/hybrid_mesh_generator.py:64-72)."""
    t = np.asarray(tets, dtype=np.int64)
    n0, n1, n2, n3 = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
    faces = np.stack([np.stack([n0, n2, n1], 1), np.stack([n0, n1, n3], 1),
                      np.stack([n1, n2, n3], 1), np.stack([n0, n3, n2], 1)], axis=1)
    return faces


# the six faces of a hex (z-, z+, y-, y+, x-, x+) as corner positions
# n0..n7, each listed counter-clockwise seen from outside
# (mesh/hybrid_mesh_generator.py:76-86)
_HEX_QUADS = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4],
                       [2, 3, 7, 6], [0, 4, 7, 3], [1, 2, 6, 5]])


def _split_quads(q: np.ndarray):
    """Each quad (…, 4) as two triangles through its smallest node
    (mesh/hybrid_mesh_generator.py:14-27): (a,b,c),(a,c,d) when the
    smallest is a or c, else (a,b,d),(b,c,d).  Returns (…, 2, 3)."""
    a, b, c, d = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    m = np.minimum(np.minimum(a, b), np.minimum(c, d))
    ac = (m == a) | (m == c)
    t0 = np.stack([a, b, np.where(ac, c, d)], axis=-1)
    t1 = np.stack([np.where(ac, a, b), c, d], axis=-1)
    return np.stack([t0, t1], axis=-2)


def hybrid_hex_tet_arrays(Lx=2.0, Ly=0.1, Lz=0.1, nx=20, ny=8, nz=8, tet_fraction=0.5,
                          boundary_names=None):
    """Arrays of ``generate_hybrid_hex_tet_mesh``
    (mesh/hybrid_mesh_generator.py:93-270), vectorised: hexes on the inlet
    side (the last hex layer splits its x+ quad into the two triangles the
    tets use), and beyond ``nx_hex = max(1, int(nx (1 - tet_fraction)))``
    each hex cut into 12 tets around an inserted centre node (two per quad
    face).  Same node numbering, cell order (k, j, i; a hex region cell's 12
    tets in face and triangle order), face lists and boundary faces (in the
    order the reference meets them) as the reference.

    Returns (nodes, RaggedCells, boundary_faces_dict, cell_zone_map); the
    boundary faces of a patch are (m, 4) rows, triangles padded with -1."""
    if boundary_names is None:
        boundary_names = {"x_min": "inlet", "x_max": "outlet", "y_min": "wall_bottom",
                          "y_max": "wall_top", "z_min": "wall_front", "z_max": "wall_back"}
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    nxh = max(1, int(nx * (1.0 - tet_fraction)))
    nxt = nx - nxh
    nbase = (nx + 1) * (ny + 1) * (nz + 1)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    base = np.stack([ii.ravel() * dx, jj.ravel() * dy, kk.ravel() * dz], axis=1)
    kc, jc, ic = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nxh, nx), indexing="ij")
    cen = np.stack([(ic.ravel() + 0.5) * dx, (jc.ravel() + 0.5) * dy, (kc.ravel() + 0.5) * dz], axis=1)
    nodes = np.concatenate([base, cen.reshape(-1, 3)])

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    corners = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                        nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                        nid(i, j + 1, k + 1)], axis=-1).astype(np.int64)      # (nz, ny, nx, 8)
    quads = corners[..., _HEX_QUADS]                                           # (nz, ny, nx, 6, 4)
    per_row = nxh + 12 * nxt
    # cell nodes, padded to 8
    cnodes = np.full((nz, ny, per_row, 8), -1, np.int64)
    cnodes[:, :, :nxh] = corners[:, :, :nxh]
    # faces, padded to 7 per cell and 4 per face
    cfaces = np.full((nz, ny, per_row, 7, 4), -1, np.int64)
    cfaces[:, :, :nxh, :6] = quads[:, :, :nxh]
    if nxt > 0:
        cfaces[:, :, nxh - 1, 6] = -1
        tri_if = _split_quads(quads[:, :, nxh - 1, 5])                        # (nz, ny, 2, 3)
        cfaces[:, :, nxh - 1, 5, :] = -1
        cfaces[:, :, nxh - 1, 5, :3] = tri_if[:, :, 0]
        cfaces[:, :, nxh - 1, 6, :3] = tri_if[:, :, 1]
        ctr = nbase + np.arange(nz * ny * nxt, dtype=np.int64).reshape(nz, ny, nxt)
        tris = _split_quads(quads[:, :, nxh:])                                # (nz, ny, nxt, 6, 2, 3)
        a, b, d = tris[..., 0], tris[..., 1], tris[..., 2]
        c = np.broadcast_to(ctr[:, :, :, None, None], a.shape)
        tets = np.stack([a, b, d, c], axis=-1).reshape(nz, ny, 12 * nxt, 4)
        cnodes[:, :, nxh:, :4] = tets
        tf = np.stack([np.stack([a, d, b], -1), np.stack([a, b, c], -1),
                       np.stack([b, d, c], -1), np.stack([a, c, d], -1)], axis=-2)
        cfaces[:, :, nxh:, :4, :3] = tf.reshape(nz, ny, 12 * nxt, 4, 3)
    cnodes = cnodes.reshape(-1, 8)
    cfaces = cfaces.reshape(-1, 7, 4)
    valid = cfaces[:, :, 0] >= 0
    faces = cfaces[valid]
    ptr = np.concatenate([[0], np.cumsum(valid.sum(axis=1))]).astype(np.int64)
    is_hex = np.zeros((nz, ny, per_row), bool)
    is_hex[:, :, :nxh] = True
    is_hex = is_hex.ravel()
    zones = {"hex": np.nonzero(is_hex)[0], "tet": np.nonzero(~is_hex)[0]}
    # boundary faces: met once, no centre node, all nodes on one outer plane
    keys = _face_keys(faces)
    face_of_entry, _, _ = _first_encounter_faces(keys)
    once = np.bincount(face_of_entry)[face_of_entry] == 1
    live = faces >= 0
    cand = once & ~np.any(live & (faces >= nbase), axis=1)
    fidx = np.nonzero(cand)[0]
    fl = faces[fidx]
    lv = fl >= 0
    xyz = nodes[np.where(lv, fl, 0)]                                          # (m, 4, 3)
    plane = np.full(len(fidx), -1, np.int64)
    for pi, (axis, val) in enumerate([(0, 0.0), (0, Lx), (1, 0.0), (1, Ly), (2, 0.0), (2, Lz)]):
        on = np.all(~lv | (np.abs(xyz[:, :, axis] - val) < 1e-10), axis=1)
        plane = np.where((plane < 0) & on, pi, plane)
    keys6 = ["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"]
    plane_name = [boundary_names.get(kname, kname) for kname in keys6]
    bdict = {boundary_names[kname]: np.empty((0, 4), np.int64) for kname in boundary_names}
    for name in dict.fromkeys(plane_name):
        mine = [pi for pi in range(6) if plane_name[pi] == name]
        bdict[name] = fl[np.isin(plane, mine)]        # in the order the faces are met
    return nodes, RaggedCells(cnodes, faces, ptr), bdict, zones


def generate_hybrid_hex_tet_mesh(Lx=2.0, Ly=0.1, Lz=0.1, nx=20, ny=8, nz=8, tet_fraction=0.5,
                                 boundary_names=None) -> FVMesh:
    """Drop-in ``generate_hybrid_hex_tet_mesh``
    (mesh/hybrid_mesh_generator.py:93-270): same arguments, same mesh."""
    nodes, rag, bdict, zones = hybrid_hex_tet_arrays(Lx, Ly, Lz, nx, ny, nz, tet_fraction,
                                                     boundary_names)
    return build_fvmesh_from_arrays(nodes, None, bdict, cell_zone_map=zones, ndim=3,
                                    cell_face_node_list=rag)
