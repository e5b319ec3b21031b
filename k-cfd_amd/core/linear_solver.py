"""Drop-in ``solve_linear_system`` (core/linear_solver.py:14-70).

Same signature and defaults.  ``method='bicgstab'`` runs the device
BiCGSTAB (libkcfd) with scipy's semantics: the reference calls
``bicgstab(A, b, x0=x0, atol=tol, maxiter=maxiter, M=M)`` and leaves scipy's
default ``rtol=1e-5`` (linear_solver.py:59), so the drop-in stops at
``||r|| < max(tol, rtol*||b||)`` with ``rtol`` defaulting to 1e-5 unless an
``rtol=`` keyword is passed.  Extra keywords: ``rtol``, ``precision``
('fp64' | 'mixed'), ``inner_rtol`` (mixed: residual-replacement trigger,
default 0.1), ``order`` ('natural' | 'color' | 'segment', the ILU(0) row ordering),
``return_info``.  A non-converged solve
prints the reference's message and returns x; a breakdown returns x
silently (linear_solver.py:60-62).

``backend``: 'auto' and 'gpu' both use the device solver (the reference's
'auto' never tried its GPU branch; there is one implementation here).
'cpu' is not provided by this package.  ``method='direct'`` is the
reference's scipy ``spsolve`` call, kept as is: it is not part of the GPU
path.

``method='gmres'`` runs the device GMRES (KCFD_METHOD_GMRES, FP64) with the
reference's call ``gmres(A, b, x0=x0, atol=tol, maxiter=maxiter, M=M)``
(linear_solver.py:64-68): scipy's restart 20 (``restart=`` keyword to
change it), ``maxiter`` counting restart cycles, rtol 1e-5 unless given, and
the reference's message on a non-converged solve.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import spsolve

from .. import _lib as L
from ..device import DeviceMatrix, default_opts
from .fvm_operators import FVMSystem

_PCS = ("none", "jacobi", "ilu0")


def _device_system(system):
    """(DeviceMatrix, b device pointer, keepalive) for an FVMSystem or CSR."""
    ctx = L.default_context()
    if isinstance(system, FVMSystem) and system.device_matrix is not None and not system.has_host_entries():
        A = system.device_matrix
        return A, A.rhs_ptr, None
    A = DeviceMatrix.from_csr(ctx, system.to_sparse())
    b = L.DevBuf.from_array(ctx, system.rhs)
    return A, b.p, b


def solve_linear_system(system: FVMSystem, x0: np.ndarray = None, method: str = "direct",
                        tol: float = 1e-8, maxiter: int = 1000, backend: str = "auto",
                        preconditioner: str = "none", **precond_kwargs) -> np.ndarray:
    if backend not in ("auto", "gpu", "cpu"):
        raise ValueError(f"Unknown backend: {backend}")
    if backend == "cpu":
        raise NotImplementedError("this package runs the solve on the GPU only (backend 'auto'/'gpu')")
    if method == "direct":
        return spsolve(system.to_sparse(), system.rhs)
    if preconditioner in ("amg",):
        print("  [Preconditioner] pyamg 미설치. ILU(0)으로 폴백합니다.")
        preconditioner = "ilu0"
    if preconditioner == "iluk":
        raise NotImplementedError("iluk has no GPU implementation; use 'ilu0'")
    if preconditioner not in _PCS:
        raise ValueError(f"Unknown preconditioner: {preconditioner}")
    if method not in ("bicgstab", "gmres"):
        raise ValueError(f"Unknown solver method: {method}")
    rtol = float(precond_kwargs.get("rtol", 1e-5))
    precision = precond_kwargs.get("precision", "fp64")
    if method == "gmres" and precision != "fp64":
        raise ValueError("gmres runs in FP64 only")
    return_info = bool(precond_kwargs.get("return_info", False))
    A, b_ptr, keep = _device_system(system)
    ctx = A.ctx
    x = L.DevBuf.from_array(ctx, np.zeros(system.n) if x0 is None else np.asarray(x0, dtype=np.float64))
    opts = default_opts(rtol=rtol, atol=tol, maxiter=maxiter, precond=preconditioner, precision=precision,
                        inner_rtol=float(precond_kwargs.get("inner_rtol", 0.1)), method=method,
                        restart=int(precond_kwargs.get("restart", 0) or 0))
    # preconditioner from the current values (as the reference rebuilds it on
    # every call, core/linear_solver.py:56); order='color' selects the
    # coloured ILU(0) row ordering
    A.setup_precond(preconditioner, precision, order=precond_kwargs.get("order"))
    res = A.solve_dev(b_ptr, x.p, opts, setup=False)
    out = x.download()
    if res.info > 0:
        print(f"[{method}] 수렴 실패: {res.info} 반복 후 중단")
    if return_info:
        return out, res.as_dict()
    return out


def solve_pressure_correction(system: FVMSystem, method: str = "direct", tol: float = 1e-6,
                              maxiter: int = 2000, preconditioner: str = "none", **precond_kwargs) -> np.ndarray:
    """Drop-in ``solve_pressure_correction`` (core/linear_solver.py:73-93).

    'direct' is scipy spsolve as in the reference.  Any other ``method`` is
    the reference's CG branch, ``cg(A, b, x0=0, atol=tol, maxiter=maxiter,
    M=M)`` with scipy's default ``rtol=1e-5``: here the device CG in FP64
    (libkcfd, KCFD_METHOD_CG) with scipy's recurrences, preconditioned by
    'none', 'jacobi' or 'ilu0' ('amg' falls back to 'ilu0' with the
    reference's message).  Extra keywords: ``rtol``, ``order`` (ILU(0) row
    ordering), ``return_info``.  As in the reference, a non-converged solve
    returns x without a message."""
    if method == "direct":
        return spsolve(system.to_sparse(), system.rhs)
    if preconditioner == "amg":
        print("  [Preconditioner] pyamg 미설치. ILU(0)으로 폴백합니다.")
        preconditioner = "ilu0"
    if preconditioner == "iluk":
        raise NotImplementedError("iluk has no GPU implementation; use 'ilu0'")
    if preconditioner not in _PCS:
        raise ValueError(f"Unknown preconditioner: {preconditioner}")
    rtol = float(precond_kwargs.get("rtol", 1e-5))
    return_info = bool(precond_kwargs.get("return_info", False))
    A, b_ptr, keep = _device_system(system)
    ctx = A.ctx
    x = L.DevBuf.from_array(ctx, np.zeros(system.n))
    opts = default_opts(rtol=rtol, atol=tol, maxiter=maxiter, precond=preconditioner, precision="fp64",
                        method="cg")
    A.setup_precond(preconditioner, "fp64", order=precond_kwargs.get("order"))
    res = A.solve_dev(b_ptr, x.p, opts, setup=False)
    out = x.download()
    if return_info:
        return out, res.as_dict()
    return out
