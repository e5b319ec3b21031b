"""Replacement of the CuPy module core/gpu_solver.py.

``detect_gpu_backend`` (:26-52) reports 'hip' when libkcfd loads and a
device is visible ('cpu' otherwise, with the reference's install hint);
``gpu_spmv`` (:55-79) and ``gpu_bicgstab`` (:82-108) run on the device.
``gpu_bicgstab`` keeps the reference's meaning: unpreconditioned, stop at
||r||/||b|| < tol.  There is no CPU fallback: errors raise KcfdError.
``benchmark_solvers`` (:202-250) times the same three solvers as the
reference and returns the same keys; its CPU BiCGSTAB is scipy's
(unpreconditioned, the same relative tolerance) in place of the reference's
private numpy loop, and it only measures.
"""

from __future__ import annotations

import time

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


def benchmark_solvers(A_csr, b: np.ndarray, n_runs: int = 3) -> dict:
    """CPU direct vs CPU BiCGSTAB vs GPU BiCGSTAB (core/gpu_solver.py:202-250):
    median wall seconds of ``n_runs`` solves each, tol 1e-6, at most 1000
    iterations, x0 = 0; the GPU solver is warmed up once first.  Returns
    {'backend', 'cpu_direct', 'x_direct', 'cpu_bicgstab', 'gpu_bicgstab'}."""
    from scipy.sparse.linalg import bicgstab, spsolve
    A = sparse.csr_matrix(A_csr)
    b = np.asarray(b, dtype=np.float64)
    results = {"backend": detect_gpu_backend()}
    times = []
    for _ in range(n_runs):
        t0 = time.time()
        x_direct = spsolve(A, b)
        times.append(time.time() - t0)
    results["cpu_direct"] = float(np.median(times))
    results["x_direct"] = x_direct
    times = []
    for _ in range(n_runs):
        t0 = time.time()
        bicgstab(A, b, rtol=1e-6, atol=0.0, maxiter=1000)
        times.append(time.time() - t0)
    results["cpu_bicgstab"] = float(np.median(times))
    if results["backend"] != "hip":
        results["gpu_bicgstab"] = float("nan")
        return results
    gpu_bicgstab(A, b, None, 1e-6, 1000)          # warm-up
    times = []
    for _ in range(n_runs):
        t0 = time.time()
        gpu_bicgstab(A, b, None, 1e-6, 1000)
        times.append(time.time() - t0)
    results["gpu_bicgstab"] = float(np.median(times))
    return results
