"""Replacement of the CuPy module core/gpu_solver.py.

``detect_gpu_backend`` (:26-52) reports 'hip' when libkcfd loads and a
device is visible ('cpu' otherwise, with the reference's install hint);
``gpu_spmv`` (:55-79) and ``gpu_bicgstab`` (:82-108) run on the device.
``gpu_bicgstab`` keeps the reference's meaning: unpreconditioned, stop at
||r||/||b|| < tol.  There is no CPU fallback: errors raise KcfdError.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse

from .. import _lib as L
from ..device import DeviceMatrix


def detect_gpu_backend() -> str:
    try:
        L.default_context()
        return "hip"
    except Exception:
        print("  [GPU] libkcfd / HIP device not available.")
        return "cpu"


def gpu_spmv(A_csr, x: np.ndarray) -> np.ndarray:
    A = DeviceMatrix.from_csr(L.default_context(), sparse.csr_matrix(A_csr))
    xd = L.DevBuf.from_array(A.ctx, x)
    yd = L.DevBuf(A.ctx, A.n)
    A.spmv(xd, yd)
    return yd.download()


def gpu_bicgstab(A_csr, b: np.ndarray, x0: np.ndarray = None, tol: float = 1e-6,
                 max_iter: int = 1000) -> np.ndarray:
    A = DeviceMatrix.from_csr(L.default_context(), sparse.csr_matrix(A_csr))
    x, _ = A.solve(b, x0, rtol=tol, atol=0.0, maxiter=max_iter, precond="none")
    return x
