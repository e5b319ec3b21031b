"""The reference's only numeric pin for the solve: case 16.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``case16_system`` restates the inline convection-diffusion assembly of
verification/case16_preconditioner.py:23-97 on the 80x80 unit channel
(:121-127): gamma = 0.01, u = 1, F = u*A*n_x; Dirichlet inlet 1 and walls 0,
zero-gradient outlet.  The reference's committed figure
figures/case16_preconditioner.png prints the BiCGSTAB iteration counts for
``atol=1e-10, rtol=0, maxiter=2000, x0=0`` (:170-173): none 105, jacobi 95,
ilu0 (spilu) 68.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse

from . import literal

CASE16_ITERS = {"none": 105, "jacobi": 95, "ilu0": 68}


def channel_arrays(L, H, nx, ny):
    """Node/cell/patch arrays of generate_channel_mesh (mesh/mesh_generator.py:96-110)."""
    dx, dy = L / nx, H / ny
    nodes = np.zeros(((nx + 1) * (ny + 1), 2))
    for j in range(ny + 1):
        for i in range(nx + 1):
            nodes[j * (nx + 1) + i] = [i * dx, j * dy]
    cells = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            cells.append(np.array([a, a + 1, a + 1 + nx + 1, a + nx + 1]))
    bf = {"wall_bottom": [np.array([i, i + 1]) for i in range(nx)],
          "wall_top": [np.array([ny * (nx + 1) + i, ny * (nx + 1) + i + 1]) for i in range(nx)],
          "inlet": [np.array([j * (nx + 1), (j + 1) * (nx + 1)]) for j in range(ny)],
          "outlet": [np.array([j * (nx + 1) + nx, (j + 1) * (nx + 1) + nx]) for j in range(ny)]}
    return nodes, cells, bf


def case16_system(nx=80, ny=80, gamma=0.01, u=1.0):
    nodes, cells, bf = channel_arrays(1.0, 1.0, nx, ny)
    m = literal.build_mesh(nodes, cells, bf, ndim=2)
    n = m.n_cells
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    dval = {"inlet": 1.0, "wall_bottom": 0.0, "wall_top": 0.0}
    dface = {}
    for name, v in dval.items():
        for fid in m.boundary_patches.get(name, []):
            dface[fid] = v
    for fid, f in enumerate(m.faces):
        o = f.owner
        F = u * f.area * f.normal[0]
        if f.neighbour >= 0:
            nb = f.neighbour
            d = float(np.linalg.norm(m.cells[nb].center - m.cells[o].center))
            if d < 1e-30:
                continue
            c = gamma * f.area / d
            rows += [o, nb, o, nb]; cols += [o, nb, nb, o]; vals += [c, c, -c, -c]
            rows += [o, o, nb, nb]; cols += [o, nb, nb, o]
            vals += [max(F, 0.0), min(F, 0.0), max(-F, 0.0), min(-F, 0.0)]
        elif fid in dface:
            d = float(np.linalg.norm(f.center - m.cells[o].center))
            if d < 1e-30:
                continue
            c = gamma * f.area / d
            rows.append(o); cols.append(o); vals.append(c)
            rhs[o] += c * dface[fid]
            if F >= 0:
                rows.append(o); cols.append(o); vals.append(F)
            else:
                rhs[o] += -F * dface[fid]
        elif F >= 0:
            rows.append(o); cols.append(o); vals.append(F)
    A = sparse.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    return A, rhs
