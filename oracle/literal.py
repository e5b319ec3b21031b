"""Literal loop transcriptions of the reference hot-path functions.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  These follow the
reference's per-face / per-cell Python loops one step at a time so that the
vectorised restatement (oracle/fvm.py) and the vectorised mesh builder of the
product package can be checked against them on tiny meshes.  They are far
too slow for anything above a few thousand cells.

Each function cites the reference lines it restates.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse


# ---------------------------------------------------------------------------
# mesh (mesh/mesh_reader.py)
# ---------------------------------------------------------------------------

class LFace:
    def __init__(self, nodes, owner, area, normal, center):
        self.nodes = nodes
        self.owner = owner
        self.neighbour = -1
        self.area = area
        self.normal = normal
        self.center = center


class LCell:
    def __init__(self, nodes, center, volume=0.0):
        self.nodes = nodes
        self.center = center
        self.volume = volume
        self.faces = []


class LMesh:
    def __init__(self, ndim):
        self.ndim = ndim
        self.cells = []
        self.faces = []
        self.boundary_patches = {}
        self.n_cells = 0
        self.n_faces = 0


def _face_geom_2d(nodes, fn):
    """mesh_reader.py:66-78."""
    a = nodes[fn[0]]
    b = nodes[fn[1]]
    c = 0.5 * (a + b)
    ex = b[0] - a[0]
    ey = b[1] - a[1]
    ln = np.sqrt(ex ** 2 + ey ** 2)
    nv = np.array([ey, -ex])
    if ln > 1e-15:
        nv /= ln
    return c, nv, ln


def _face_geom_3d(nodes, fn):
    """Newell normal, mesh_reader.py:81-99."""
    p = nodes[fn]
    m = len(p)
    c = np.mean(p, axis=0)
    nv = np.zeros(3)
    for a in range(m):
        b = (a + 1) % m
        nv[0] += (p[a, 1] - p[b, 1]) * (p[a, 2] + p[b, 2])
        nv[1] += (p[a, 2] - p[b, 2]) * (p[a, 0] + p[b, 0])
        nv[2] += (p[a, 0] - p[b, 0]) * (p[a, 1] + p[b, 1])
    ar = np.linalg.norm(nv)
    if ar > 1e-15:
        nv /= ar
    return c, nv, ar * 0.5


def build_mesh(node_coords, cell_node_list, boundary_faces_dict, ndim=2,
               cell_face_node_list=None) -> LMesh:
    """mesh_reader.py:141-323, one face at a time."""
    nodes = np.array(node_coords, dtype=np.float64)
    m = LMesh(ndim)
    if ndim == 3 and cell_face_node_list is not None:
        for cn in cell_node_list:
            cn = np.array(cn, dtype=int)
            m.cells.append(LCell(cn, np.mean(nodes[cn], axis=0)))
        lookup = {}
        for ci, flist in enumerate(cell_face_node_list):
            for fn in flist:
                fn = np.array(fn, dtype=int)
                key = tuple(sorted(int(v) for v in fn))
                if key in lookup:
                    fid = lookup[key]
                    m.faces[fid].neighbour = ci
                    m.cells[ci].faces.append(fid)
                else:
                    c, nv, ar = _face_geom_3d(nodes, fn)
                    if np.dot(nv, c - m.cells[ci].center) < 0:
                        nv = -nv
                    m.faces.append(LFace(fn, ci, ar, nv, c))
                    lookup[key] = len(m.faces) - 1
                    m.cells[ci].faces.append(len(m.faces) - 1)
        for cell in m.cells:   # mesh_reader.py:116-138
            v = 0.0
            for fid in cell.faces:
                f = m.faces[fid]
                sg = 1.0
                if f.owner != -1 and np.dot(f.normal, f.center - cell.center) < 0:
                    sg = -1.0
                v += sg * np.dot(f.center, f.normal) * f.area
            v = abs(v) / 3.0
            cell.volume = v if v >= 1e-30 else 1e-30
        keyf = lambda fn: tuple(sorted(int(v) for v in fn))
    else:
        for cn in cell_node_list:
            cn = np.array(cn, dtype=int)
            p = nodes[cn]
            k = len(p)
            sh = 0.0
            for a in range(k):
                b = (a + 1) % k
                sh += p[a, 0] * p[b, 1] - p[b, 0] * p[a, 1]
            m.cells.append(LCell(cn, np.mean(p, axis=0), abs(sh) / 2.0))
        lookup = {}
        for ci, cn in enumerate(cell_node_list):
            cn = np.array(cn, dtype=int)
            k = len(cn)
            for a in range(k):
                n0, n1 = int(cn[a]), int(cn[(a + 1) % k])
                key = (min(n0, n1), max(n0, n1))
                if key in lookup:
                    fid = lookup[key]
                    m.faces[fid].neighbour = ci
                    m.cells[ci].faces.append(fid)
                else:
                    fn = np.array([n0, n1], dtype=int)
                    c, nv, ln = _face_geom_2d(nodes, fn)
                    if np.dot(nv, c - m.cells[ci].center) < 0:
                        nv = -nv
                    m.faces.append(LFace(fn, ci, ln, nv, c))
                    lookup[key] = len(m.faces) - 1
                    m.cells[ci].faces.append(len(m.faces) - 1)
        keyf = lambda fn: (min(int(fn[0]), int(fn[1])), max(int(fn[0]), int(fn[1])))
    tag = {}
    for name, flist in boundary_faces_dict.items():
        for fn in flist:
            tag[keyf(np.array(fn, dtype=int))] = name
    for fid, f in enumerate(m.faces):
        if f.neighbour == -1:
            name = tag.get(keyf(f.nodes), "default")
            m.boundary_patches.setdefault(name, []).append(fid)
    m.n_cells = len(m.cells)
    m.n_faces = len(m.faces)
    return m


# ---------------------------------------------------------------------------
# hybrid hex/tet generator (mesh/hybrid_mesh_generator.py)
# ---------------------------------------------------------------------------

def _split_quad(q):
    """hybrid_mesh_generator.py:14-27: two triangles through the quad's
    smallest node, so that both cells sharing the quad split it alike."""
    a, b, c, d = (int(v) for v in q)
    lo = min(a, b, c, d)
    return [(a, b, c), (a, c, d)] if lo in (a, c) else [(a, b, d), (b, c, d)]


def _hex_quads(h):
    """hybrid_mesh_generator.py:76-86 (and :50-57): z-, z+, y-, y+, x-, x+."""
    return [[h[0], h[3], h[2], h[1]], [h[4], h[5], h[6], h[7]], [h[0], h[1], h[5], h[4]],
            [h[2], h[3], h[7], h[6]], [h[0], h[4], h[7], h[3]], [h[1], h[2], h[6], h[5]]]


def hybrid_hex_tet(Lx=2.0, Ly=0.1, Lz=0.1, nx=20, ny=8, nz=8, tet_fraction=0.5,
                   boundary_names=None):
    """generate_hybrid_hex_tet_mesh (hybrid_mesh_generator.py:93-270) up to
    the builder call, one cell at a time: returns (nodes, cell_node_list,
    cell_face_node_list, boundary_faces_dict, cell_zone_map)."""
    if boundary_names is None:
        boundary_names = {"x_min": "inlet", "x_max": "outlet", "y_min": "wall_bottom",
                          "y_max": "wall_top", "z_min": "wall_front", "z_max": "wall_back"}
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    nxh = max(1, int(nx * (1.0 - tet_fraction)))
    nxt = nx - nxh
    nbase = (nx + 1) * (ny + 1) * (nz + 1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    nodes = np.zeros((nbase + nxt * ny * nz, 3))
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                nodes[nid(i, j, k)] = [i * dx, j * dy, k * dz]
    centre = {}
    nxt_id = nbase
    for k in range(nz):
        for j in range(ny):
            for i in range(nxh, nx):
                nodes[nxt_id] = [(i + 0.5) * dx, (j + 0.5) * dy, (k + 0.5) * dz]
                centre[(i, j, k)] = nxt_id
                nxt_id += 1
    planes = {}
    for key, axis, val in [("x_min", 0, 0.0), ("x_max", 0, Lx), ("y_min", 1, 0.0), ("y_max", 1, Ly),
                           ("z_min", 2, 0.0), ("z_max", 2, Lz)]:
        planes[key] = {v for v in range(nbase) if abs(nodes[v, axis] - val) < 1e-10}
    cells, cell_faces, zones = [], [], {"hex": [], "tet": []}
    count = {}
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                h = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)]
                if i < nxh:
                    fl = [np.array(q, dtype=int) for q in _hex_quads(h)]
                    if i == nxh - 1 and nxt > 0:   # interface: the x+ quad as the tets split it
                        fl = fl[:5] + [np.array(t, dtype=int) for t in _split_quad(fl[5])]
                    zones["hex"].append(len(cells))
                    cells.append(np.array(h, dtype=int))
                    cell_faces.append(fl)
                else:
                    c = centre[(i, j, k)]
                    for q in _hex_quads(h):
                        for a, b, d in _split_quad(q):
                            zones["tet"].append(len(cells))
                            cells.append(np.array([a, b, d, c], dtype=int))
                            cell_faces.append([np.array(f, dtype=int) for f in
                                               ([a, d, b], [a, b, c], [b, d, c], [a, c, d])])
    for fl in cell_faces:
        for f in fl:
            key = tuple(sorted(int(v) for v in f))
            count[key] = count.get(key, 0) + 1
    bdict = {boundary_names[k]: [] for k in boundary_names}
    seen = set()
    for fl in cell_faces:
        for f in fl:
            key = tuple(sorted(int(v) for v in f))
            if key in seen or count[key] != 1 or any(int(v) >= nbase for v in f):
                continue
            seen.add(key)
            for pk, pset in planes.items():
                if set(int(v) for v in f) <= pset:
                    bdict.setdefault(boundary_names.get(pk, pk), []).append(f)
                    break
    return nodes, cells, cell_faces, bdict, zones


# ---------------------------------------------------------------------------
# FVMSystem and operators (core/fvm_operators.py)
# ---------------------------------------------------------------------------

class LSystem:
    """fvm_operators.py:13-56."""

    def __init__(self, n):
        self.n = n
        self.rows, self.cols, self.vals = [], [], []
        self.rhs = np.zeros(n)
        self.diag = np.zeros(n)

    def add_diagonal(self, i, v):
        self.rows.append(i); self.cols.append(i); self.vals.append(v)
        self.diag[i] += v

    def add_off_diagonal(self, i, j, v):
        self.rows.append(i); self.cols.append(j); self.vals.append(v)

    def add_source(self, i, v):
        self.rhs[i] += v

    def to_sparse(self):
        return sparse.coo_matrix((self.vals, (self.rows, self.cols)),
                                 shape=(self.n, self.n)).tocsr()


def diffusion(m: LMesh, gamma, s: LSystem):
    """fvm_operators.py:59-100."""
    for f in m.faces:
        if f.neighbour < 0:
            continue
        o, nb = f.owner, f.neighbour
        d = np.linalg.norm(m.cells[nb].center - m.cells[o].center)
        if d < 1e-30:
            continue
        go, gn = gamma[o], gamma[nb]
        gf = 2.0 * go * gn / (go + gn) if go + gn > 1e-30 else 0.0
        c = gf * f.area / d
        s.add_diagonal(o, c)
        s.add_diagonal(nb, c)
        s.add_off_diagonal(o, nb, -c)
        s.add_off_diagonal(nb, o, -c)


def upwind(m: LMesh, F, s: LSystem):
    """fvm_operators.py:103-129."""
    for fid, f in enumerate(m.faces):
        o = f.owner
        q = F[fid]
        if f.neighbour >= 0:
            nb = f.neighbour
            s.add_diagonal(o, max(q, 0.0))
            s.add_off_diagonal(o, nb, min(q, 0.0))
            s.add_diagonal(nb, max(-q, 0.0))
            s.add_off_diagonal(nb, o, min(-q, 0.0))
        elif q >= 0:
            s.add_diagonal(o, q)


def temporal(m: LMesh, rho, dt, phi_old, s: LSystem):
    """fvm_operators.py:132-148."""
    for ci, c in enumerate(m.cells):
        k = rho * c.volume / dt
        s.add_diagonal(ci, k)
        s.add_source(ci, k * phi_old[ci])


def source(m: LMesh, sv, s: LSystem):
    """fvm_operators.py:151-161."""
    for ci, c in enumerate(m.cells):
        s.add_source(ci, sv[ci] * c.volume)


def linearized(m: LMesh, Sp, Su, s: LSystem):
    """fvm_operators.py:164-175."""
    for ci, c in enumerate(m.cells):
        s.add_diagonal(ci, -Sp[ci] * c.volume)
        s.add_source(ci, Su[ci] * c.volume)


def boundary(m: LMesh, phi_bnd: dict, gamma, F, s: LSystem, bc_types: dict):
    """fvm_operators.py:178-225."""
    for name, fids in m.boundary_patches.items():
        bc = bc_types.get(name, {"type": "zero_gradient"})
        kind = bc["type"]
        for li, fid in enumerate(fids):
            f = m.faces[fid]
            o = f.owner
            d = np.linalg.norm(f.center - m.cells[o].center)
            if d < 1e-30:
                continue
            if kind == "dirichlet":
                pb = phi_bnd[name][li]
                c = gamma[o] * f.area / d
                s.add_diagonal(o, c)
                s.add_source(o, c * pb)
                if F[fid] < 0:
                    s.add_source(o, -F[fid] * pb)
            elif kind == "neumann":
                s.add_source(o, bc.get("value", 0.0) * f.area)
            elif kind == "zero_gradient":
                if F[fid] < 0:
                    s.add_diagonal(o, -F[fid])


def under_relax(s: LSystem, phi, alpha):
    """fvm_operators.py:228-240."""
    if alpha >= 1.0:
        return
    for ci in range(s.n):
        ap = s.diag[ci]
        s.add_diagonal(ci, ap * (1.0 - alpha) / alpha)
        s.add_source(ci, ap * (1.0 - alpha) / alpha * phi[ci])


# ---------------------------------------------------------------------------
# gradient / interpolation (core/gradient.py, core/interpolation.py)
# ---------------------------------------------------------------------------

def _weight(m: LMesh, fid):
    """gradient.py:102-114 (identical to interpolation.py:185-196)."""
    f = m.faces[fid]
    do = np.linalg.norm(f.center - m.cells[f.owner].center)
    dn = np.linalg.norm(f.center - m.cells[f.neighbour].center)
    t = do + dn
    return 0.5 if t < 1e-30 else dn / t


def _bvalue(m: LMesh, values, bvals, fid):
    """gradient.py:117-127: linear search over the patches."""
    for name, fids in m.boundary_patches.items():
        if fid in fids:
            return bvals[name][list(fids).index(fid)]
    return values[m.faces[fid].owner]


def face_value(m: LMesh, values, bvals, fid, mode, flux_dir=0.0, grad=None):
    """interpolation.py:55-134 for one face: upwind_face_value,
    central_face_value, tvd_face_value."""
    f = m.faces[fid]
    if f.neighbour < 0:
        return _bvalue(m, values, bvals, fid)
    if mode == "central":
        w = _weight(m, fid)
        return w * values[f.owner] + (1.0 - w) * values[f.neighbour]
    up, dn = (f.owner, f.neighbour) if flux_dir >= 0 else (f.neighbour, f.owner)
    if mode == "upwind":
        return values[up]
    pu, pd = values[up], values[dn]
    dx = f.center[:m.ndim] - m.cells[up].center[:m.ndim]
    lin = pu + np.dot(grad[up], dx)
    dup, dtot = lin - pu, pd - pu
    if abs(dtot) < 1e-15:
        return pu
    r = 2.0 * dup / dtot - 1.0
    psi = (r + abs(r)) / (1.0 + abs(r)) if (1.0 + abs(r)) > 1e-15 else 0.0
    return pu + 0.5 * psi * dtot


def green_gauss(m: LMesh, values, bvals):
    """gradient.py:12-48."""
    nd = m.ndim
    g = np.zeros((m.n_cells, nd))
    for fid, f in enumerate(m.faces):
        if f.neighbour >= 0:
            w = _weight(m, fid)
            pf = w * values[f.owner] + (1.0 - w) * values[f.neighbour]
        else:
            pf = _bvalue(m, values, bvals, fid)
        fl = pf * f.normal[:nd] * f.area
        g[f.owner] += fl
        if f.neighbour >= 0:
            g[f.neighbour] -= fl
    for ci, c in enumerate(m.cells):
        if c.volume > 1e-30:
            g[ci] /= c.volume
    return g


def least_squares(m: LMesh, values, bvals):
    """gradient.py:51-99."""
    nd = m.ndim
    g = np.zeros((m.n_cells, nd))
    for ci, c in enumerate(m.cells):
        xp = c.center[:nd]
        dxs, dphis = [], []
        for fid in c.faces:
            f = m.faces[fid]
            if f.neighbour >= 0:
                nbr = f.neighbour if f.owner == ci else f.owner
                dxs.append(m.cells[nbr].center[:nd] - xp)
                dphis.append(values[nbr] - values[ci])
            else:
                dxs.append(f.center[:nd] - xp)
                dphis.append(_bvalue(m, values, bvals, fid) - values[ci])
        if len(dxs) < nd:
            continue
        A = np.array(dxs)
        b = np.array(dphis)
        try:
            g[ci] = np.linalg.solve(A.T @ A, A.T @ b)
        except np.linalg.LinAlgError:
            pass
    return g


LIMITERS = {
    "minmod": lambda r: np.maximum(0.0, np.minimum(1.0, r)),
    "superbee": lambda r: np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0))),
    "van_albada": lambda r: np.where(r > 0, (r * r + r) / (r * r + 1.0), 0.0),
    "van_leer": lambda r: np.where(r > 0, 2.0 * r / (1.0 + r), 0.0),
}   # interpolation.py:17-44


def muscl(m: LMesh, values, F, grad, limiter="van_leer"):
    """interpolation.py:227-303."""
    lim = LIMITERS[limiter]
    corr = np.zeros(m.n_cells)
    for fid, f in enumerate(m.faces):
        if f.neighbour < 0:
            continue
        o, nb = f.owner, f.neighbour
        q = F[fid]
        up, dn = (o, nb) if q >= 0 else (nb, o)
        pu, pd = values[up], values[dn]
        dx = f.center - m.cells[up].center
        lin = pu + np.dot(grad[up], dx)
        dface = lin - pu
        dcd = pd - pu
        if abs(dcd) < 1e-15:
            continue
        r = 2.0 * dface / dcd - 1.0
        psi = float(lim(np.array([r]))[0])
        fc = abs(q) * ((pu + 0.5 * psi * dcd) - pu)
        if q >= 0:
            corr[o] -= fc
            corr[nb] += fc
        else:
            corr[nb] -= fc
            corr[o] += fc
    return corr


def mass_flux(m: LMesh, U, Ubnd, rho):
    """interpolation.py:163-182 (boundary: patch vector value, else owner)."""
    F = np.zeros(m.n_faces)
    for fid, f in enumerate(m.faces):
        if f.neighbour >= 0:
            w = _weight(m, fid)
            uf = w * U[f.owner] + (1.0 - w) * U[f.neighbour]
        else:
            uf = _bvalue(m, U, Ubnd, fid)
        F[fid] = rho * np.dot(uf, f.normal) * f.area
    return F


# ---------------------------------------------------------------------------
# two-fluid pressure correction (models/two_fluid.py:641-725)
# ---------------------------------------------------------------------------

def two_fluid_pressure_step(m: LMesh, alpha_l, alpha_g, aP_l, aP_g, rho_l, rho_g, mf_l, mf_g,
                            U_l, U_g, p, alpha_p):
    """models/two_fluid.py:641-725, one face and one cell at a time.
    Returns (p', U_l, U_g, p, A, b, g)."""
    from scipy.sparse.linalg import spsolve
    n = m.n_cells
    s = LSystem(n)
    for fid, f in enumerate(m.faces):
        o = f.owner
        if f.neighbour >= 0:
            nb = f.neighbour
            d = np.linalg.norm(m.cells[nb].center - m.cells[o].center)
            if d < 1e-30:
                continue
            alf = 0.5 * (alpha_l[o] + alpha_l[nb])
            agf = 0.5 * (alpha_g[o] + alpha_g[nb])
            aplf = 0.5 * (aP_l[o] + aP_l[nb])
            apgf = 0.5 * (aP_g[o] + aP_g[nb])
            vf = 0.5 * (m.cells[o].volume + m.cells[nb].volume)
            c = (rho_l * alf * vf / max(aplf, 1e-30) + rho_g * agf * vf / max(apgf, 1e-30)) * f.area / d
            s.add_diagonal(o, c)
            s.add_diagonal(nb, c)
            s.add_off_diagonal(o, nb, -c)
            s.add_off_diagonal(nb, o, -c)
            tm = mf_l[fid] * alpha_l[o] + mf_g[fid] * alpha_g[o]
            s.add_source(o, -tm)
            s.add_source(nb, tm)
        else:
            tm = mf_l[fid] * alpha_l[o] + mf_g[fid] * alpha_g[o]
            s.add_source(o, -tm)
    s.add_diagonal(0, 1e10)
    A = s.to_sparse()
    pp = spsolve(A.tocsc(), s.rhs)
    if np.any(np.isnan(pp)):
        pp = np.zeros(n)
    p = np.array(p, np.float64) + alpha_p * pp
    nd = m.ndim
    g = np.zeros((n, nd))
    for f in m.faces:
        if f.neighbour >= 0:
            dp = pp[f.neighbour] - pp[f.owner]
            for k in range(nd):
                fl = dp * f.normal[k] * f.area
                g[f.owner, k] += fl
                g[f.neighbour, k] -= fl
    for ci in range(n):
        if m.cells[ci].volume > 1e-30:
            g[ci] /= m.cells[ci].volume
    Ul = np.array(U_l, np.float64, copy=True)
    Ug = np.array(U_g, np.float64, copy=True)
    for ci in range(n):
        v = m.cells[ci].volume
        for k in range(nd):
            Ul[ci, k] += alpha_p * (-alpha_l[ci] * v / max(aP_l[ci], 1e-20) * g[ci, k])
            Ug[ci, k] += alpha_p * (-alpha_g[ci] * v / max(aP_g[ci], 1e-20) * g[ci, k])
    return pp, Ul, Ug, p, A, s.rhs.copy(), g
