"""Face values, limiters, mass flux and MUSCL/TVD deferred correction on
the device.

Mirrors core/interpolation.py:
- the four TVD limiters (:17-44) as callables on arrays or scalars,
  ``LIMITER_REGISTRY`` and ``get_limiter`` (:38-52, the same ``ValueError``
  for an unknown name);
- the face values ``upwind_face_value``, ``central_face_value``,
  ``tvd_face_value`` and ``compute_face_values_upwind`` /
  ``compute_face_values_central`` (:55-161);
- ``compute_mass_flux`` (:163-182);
- ``muscl_deferred_correction`` (:227-303), including the reference's
  r = 2 (phi_lin - phi_U) / (phi_D - phi_U) - 1 with phi_lin taken at
  x_f - x_U.

Every value is computed by libkcfd kernels; the per-face functions evaluate
all faces on the device and return the one asked for, as the array functions
are the efficient form.
"""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from ._dev import Scratch, device_mesh


def _limit(code: int, r):
    a = np.asarray(r, dtype=np.float64)
    flat = np.ascontiguousarray(a.reshape(-1))
    ctx = L.default_context()
    with Scratch(ctx) as s:
        rd = s.put(flat)
        out = s.empty(max(flat.size, 1))
        L.check(L.lib().kcfd_limiter(ctx.h, code, rd.p, out.p, flat.size))
        psi = out.download()[:flat.size].reshape(a.shape)
    return psi[()] if a.ndim == 0 else psi


def limiter_minmod(r):
    """interpolation.py:17-19: max(0, min(1, r))."""
    return _limit(L.LIMITERS["minmod"], r)


def limiter_superbee(r):
    """interpolation.py:22-25: max(0, min(2r, 1), min(r, 2))."""
    return _limit(L.LIMITERS["superbee"], r)


def limiter_van_albada(r):
    """interpolation.py:28-30: (r^2 + r) / (r^2 + 1) where r > 0, else 0."""
    return _limit(L.LIMITERS["van_albada"], r)


def limiter_van_leer(r):
    """interpolation.py:33-35: 2r / (1 + r) where r > 0, else 0."""
    return _limit(L.LIMITERS["van_leer"], r)


LIMITER_REGISTRY = {
    "minmod": limiter_minmod,
    "superbee": limiter_superbee,
    "van_albada": limiter_van_albada,
    "van_leer": limiter_van_leer,
}


def get_limiter(name: str):
    """interpolation.py:47-52: the limiter function, or ValueError."""
    if name not in LIMITER_REGISTRY:
        raise ValueError(f"Unknown limiter: {name}. Available: {list(LIMITER_REGISTRY.keys())}")
    return LIMITER_REGISTRY[name]


def _face_values(phi, mode: int, F=None, grad=None) -> np.ndarray:
    mesh = phi.mesh
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    phib = dm.boundary_array(getattr(phi, "boundary_values", {}) or {},
                             default=lambda nm: np.asarray(phi.values)[mesh.owner[np.asarray(mesh.boundary_patches[nm])]])
    with Scratch(dm.ctx) as s:
        p, pb = s.put(phi.values), s.put(phib)
        Fd = s.put(F) if F is not None else None
        gd = s.put(np.ascontiguousarray(np.asarray(grad, dtype=np.float64)[:, :nd])) if grad is not None else None
        out = s.empty(mesh.n_faces)
        L.check(L.lib().kcfd_face_values(dm.h, mode, p.p, pb.p, Fd.p if Fd else None, gd.p if gd else None,
                                         out.p))
        return out.download()


def compute_face_values_upwind(phi, mass_flux: np.ndarray) -> np.ndarray:
    """interpolation.py:137-155: the upwind cell's value on every face (the
    owner's where the flux is >= 0), the patch value on boundary faces."""
    return _face_values(phi, 0, F=np.asarray(mass_flux, dtype=np.float64))


def compute_face_values_central(phi) -> np.ndarray:
    """interpolation.py:158-161: g phi_O + (1 - g) phi_N with the owner
    weight g = d_N / (d_O + d_N) (0.5 if the sum < 1e-30)."""
    return _face_values(phi, 1)


def _one_face_flux(phi, face_id: int, flux_dir: float) -> np.ndarray:
    F = np.zeros(phi.mesh.n_faces)
    F[face_id] = flux_dir
    return F


def upwind_face_value(phi, face_id: int, flux_dir: float) -> float:
    """interpolation.py:55-68."""
    return float(_face_values(phi, 0, F=_one_face_flux(phi, face_id, flux_dir))[face_id])


def central_face_value(phi, face_id: int) -> float:
    """interpolation.py:71-82."""
    return float(_face_values(phi, 1)[face_id])


def tvd_face_value(phi, face_id: int, flux_dir: float, grad_phi: np.ndarray) -> float:
    """interpolation.py:85-134: phi_U + 0.5 psi (phi_D - phi_U) with the van
    Leer psi = (r + |r|) / (1 + |r|), r = 2 (phi_lin - phi_U) / (phi_D -
    phi_U) - 1; phi_U when |phi_D - phi_U| < 1e-15."""
    return float(_face_values(phi, 2, F=_one_face_flux(phi, face_id, flux_dir), grad=grad_phi)[face_id])


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
    get_limiter(limiter_name)          # the reference's ValueError for an unknown name
    code = L.LIMITERS[limiter_name]
    dm = device_mesh(mesh)
    nd = getattr(mesh, "ndim", 2)
    g = np.ascontiguousarray(np.asarray(grad_phi, dtype=np.float64)[:, :nd])
    with Scratch(dm.ctx) as s:
        p, F, gd = s.put(phi.values), s.put(mass_flux), s.put(g)
        c = s.empty(mesh.n_cells)
        L.check(L.lib().kcfd_muscl(dm.h, p.p, F.p, gd.p, code, c.p))
        return c.download()
