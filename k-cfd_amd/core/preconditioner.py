"""Drop-in ``create_preconditioner`` (core/preconditioner.py:18-76).

Returns ``(M, info)`` like the reference: ``M`` is a scipy LinearOperator
whose matvec applies the preconditioner on the GPU (Jacobi
``where(|a_ii|>1e-30, 1/a_ii, 0)`` or structural ILU(0) factorised on the
device), ``info`` is ``{'method', 'setup_time'}``; 'none' gives ``(None,
...)``; an unknown name raises ``ValueError``.  As the reference does when
pyamg is absent, 'amg' prints a note and falls back to 'ilu0'.  'iluk'
(SuperLU with fill) has no GPU implementation and raises.
"""

from __future__ import annotations

import time

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import LinearOperator

from .. import _lib as L
from ..device import DeviceMatrix


class DevicePreconditioner(LinearOperator):
    """M^{-1} applied by libkcfd (kcfd_precond_apply)."""

    def __init__(self, A: DeviceMatrix, kind: str):
        super().__init__(dtype=np.float64, shape=(A.n, A.n))
        self.A = A
        self.kind = kind
        self._x = L.DevBuf(A.ctx, A.n)
        self._y = L.DevBuf(A.ctx, A.n)

    def _matvec(self, x):
        self._x.upload(np.asarray(x, dtype=np.float64).ravel())
        L.check(L.lib().kcfd_precond_apply(self.A.h, self._x.p, self._y.p))
        return self._y.download()


def create_preconditioner(A_csr, method: str = "none", **kwargs):
    t0 = time.time()
    if method == "none":
        return None, {"method": "none", "setup_time": 0.0}
    if method == "amg":
        print("  [Preconditioner] pyamg 미설치. ILU(0)으로 폴백합니다.")
        return create_preconditioner(A_csr, method="ilu0")
    if method == "iluk":
        raise NotImplementedError("iluk (SuperLU with fill) has no GPU implementation; use 'ilu0'")
    if method not in ("jacobi", "ilu0"):
        raise ValueError(f"Unknown preconditioner: {method}")
    A = A_csr if isinstance(A_csr, DeviceMatrix) else DeviceMatrix.from_csr(L.default_context(),
                                                                            sparse.csr_matrix(A_csr))
    A.setup_precond(method, "fp64", order=kwargs.get("order"))
    M = DevicePreconditioner(A, method)
    return M, {"method": method, "setup_time": time.time() - t0}
