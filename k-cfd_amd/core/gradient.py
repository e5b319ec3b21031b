"""Green-Gauss and least-squares gradients on the device (core/gradient.py).

Green-Gauss (core/gradient.py:12-48):

phi_f = g phi_O + (1-g) phi_N with g = d_N/(d_O+d_N) (0.5 if the sum
< 1e-30) on internal faces, the patch value on boundary faces; grad_O +=
phi_f n A, grad_N -= phi_f n A in face order; divided by V where V > 1e-30.
Computed by a cell-parallel gather over each cell's faces (no atomics).
"""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from ._dev import Scratch, device_mesh


def green_gauss_gradient(phi) -> np.ndarray:
    mesh = phi.mesh
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    phib = dm.boundary_array(getattr(phi, "boundary_values", {}) or {},
                             default=lambda nm: np.asarray(phi.values)[mesh.owner[np.asarray(mesh.boundary_patches[nm])]])
    with Scratch(dm.ctx) as s:
        p, pb = s.put(phi.values), s.put(phib)
        g = s.empty(mesh.n_cells * nd)
        L.check(L.lib().kcfd_grad_gg(dm.h, p.p, pb.p, g.p))
        out = g.download().reshape(mesh.n_cells, nd)
    return out


def green_gauss_gradient_device(dm, phi_dev, phib_dev, grad_dev):
    """Device-resident variant: all arguments are device buffers."""
    L.check(L.lib().kcfd_grad_gg(dm.h, phi_dev.p, phib_dev.p, grad_dev.p))


def least_squares_gradient(phi) -> np.ndarray:
    """core/gradient.py:51-99: per cell, the least-squares fit of the
    differences to its face neighbours (face centre and patch value on a
    boundary face); zero where the normal equations are singular."""
    mesh = phi.mesh
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    phib = dm.boundary_array(getattr(phi, "boundary_values", {}) or {},
                             default=lambda nm: np.asarray(phi.values)[mesh.owner[np.asarray(mesh.boundary_patches[nm])]])
    with Scratch(dm.ctx) as s:
        p, pb = s.put(phi.values), s.put(phib)
        g = s.empty(mesh.n_cells * nd)
        L.check(L.lib().kcfd_grad_lsq(dm.h, p.p, pb.p, g.p))
        out = g.download().reshape(mesh.n_cells, nd)
    return out
