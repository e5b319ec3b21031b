"""Interphase drag closure on the device (models/closure.py).

``schiller_naumann_drag(Re_p)`` (models/closure.py:13-34) and
``drag_coefficient_implicit(alpha_g, rho_l, u_g, u_l, d_b, mu_l)`` (:72-92)
with the reference's names, arguments and results; the two-fluid momentum
assembly adds ``K * V`` to the diagonal and ``K * V * u_other`` to the
right-hand side (models/two_fluid.py:592-596), which the device recipes do
with ``kcfd_asm_cell_terms``.  Every value is computed by a libkcfd kernel.
"""

from __future__ import annotations

import numpy as np

from . import _lib as L
from .core._dev import Scratch


def schiller_naumann_drag(Re_p) -> np.ndarray:
    """C_D = 24/Re (1 + 0.15 Re^0.687) for Re < 1000, else 0.44, with Re
    floored at 1e-10 (models/closure.py:13-34)."""
    a = np.asarray(Re_p, dtype=np.float64)
    flat = np.ascontiguousarray(a.reshape(-1))
    ctx = L.default_context()
    with Scratch(ctx) as s:
        rd = s.put(flat)
        out = s.empty(max(flat.size, 1))
        L.check(L.lib().kcfd_schiller_naumann(ctx.h, rd.p, out.p, flat.size))
        cd = out.download()[:flat.size].reshape(a.shape)
    return cd


def drag_coefficient_implicit(alpha_g, rho_l: float, u_g, u_l, d_b: float, mu_l: float) -> np.ndarray:
    """K = 0.75 C_D alpha_g rho_l |u_g - u_l| / d_b per cell
    (models/closure.py:72-92); u_g, u_l are (n_cells, ndim)."""
    ug = np.ascontiguousarray(u_g, dtype=np.float64)
    ul = np.ascontiguousarray(u_l, dtype=np.float64)
    if ug.ndim != 2 or ug.shape != ul.shape:
        raise ValueError("u_g and u_l must be (n_cells, ndim) arrays of one shape")
    n, nd = ug.shape
    ag = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha_g, np.float64), (n,)))
    ctx = L.default_context()
    with Scratch(ctx) as s:
        a, g, l = s.put(ag), s.put(ug.reshape(-1)), s.put(ul.reshape(-1))
        K = s.empty(max(n, 1))
        L.check(L.lib().kcfd_drag_implicit(ctx.h, n, nd, a.p, float(rho_l), g.p, l.p, float(d_b), float(mu_l),
                                           K.p))
        out = K.download()[:n]
    return out
