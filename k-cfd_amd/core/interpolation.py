"""Face mass flux and MUSCL/TVD deferred correction on the device.

Mirrors core/interpolation.py: the limiter registry (:17-52, the same four
limiters and the same ``ValueError`` for an unknown name), ``compute_mass_flux``
(:163-182) and ``muscl_deferred_correction`` (:227-303), including the
reference's r = 2 (phi_lin - phi_U)/(phi_D - phi_U) - 1 with phi_lin taken at
x_f - x_U.
"""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from ._dev import Scratch, device_mesh

LIMITER_REGISTRY = {name: code for name, code in L.LIMITERS.items()}


def get_limiter(name: str) -> int:
    """interpolation.py:47-52: the limiter code, or ValueError."""
    if name not in LIMITER_REGISTRY:
        raise ValueError(f"Unknown limiter: {name}. Available: {list(LIMITER_REGISTRY.keys())}")
    return LIMITER_REGISTRY[name]


def compute_mass_flux(U, rho: float, mesh) -> np.ndarray:
    """F_f = rho (u_f . n) A (interpolation.py:163-182)."""
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    Ub = dm.boundary_array(getattr(U, "boundary_values", {}) or {}, width=nd,
                           default=lambda nm: np.asarray(U.values)[mesh.owner[np.asarray(mesh.boundary_patches[nm])]])
    with Scratch(dm.ctx) as s:
        u, ub = s.put(U.values), s.put(Ub)
        F = s.empty(mesh.n_faces)
        L.check(L.lib().kcfd_mass_flux(dm.h, u.p, ub.p, float(rho), F.p))
        return F.download()


def muscl_deferred_correction(mesh, phi, mass_flux, grad_phi, limiter_name: str = "van_leer") -> np.ndarray:
    """Per-cell RHS correction of the MUSCL deferred-correction pattern
    (interpolation.py:227-303)."""
    code = get_limiter(limiter_name)
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    g = np.ascontiguousarray(np.asarray(grad_phi, dtype=np.float64)[:, :nd])
    with Scratch(dm.ctx) as s:
        p, F, gd = s.put(phi.values), s.put(mass_flux), s.put(g)
        c = s.empty(mesh.n_cells)
        L.check(L.lib().kcfd_muscl(dm.h, p.p, F.p, gd.p, code, c.p))
        return c.download()
