"""scipy harness with the reference's solve semantics, plus textbook ILU(0).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``solve_cg`` calls ``scipy.sparse.linalg.cg`` as ``solve_pressure_correction``
  does (core/linear_solver.py:88-93), with explicit ``rtol``/``atol`` and an
  iteration-counting callback.
* ``solve_bicgstab`` calls ``scipy.sparse.linalg.bicgstab`` exactly as
  ``solve_linear_system`` does (core/linear_solver.py:56-62) but with explicit
  ``rtol``/``atol`` and an iteration-counting callback (the convention of
  verification/case16_preconditioner.py:161-173).
* ``make_preconditioner`` restates ``create_preconditioner``
  (core/preconditioner.py:40-54): Jacobi ``where(|diag|>1e-30, 1/diag, 0)``
  and scipy ``spilu(A.tocsc(), fill_factor=1, drop_tol=0)``; ``'ilu0_struct'``
  is the textbook ILU(0) of ilu0.c that the GPU implements, ``'ilu0_color'``
  the same on P A P^T for the first-fit colour ordering.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import LinearOperator, bicgstab, cg, spilu

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "libkcfd_oracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.kcfd_oracle_ilu0_factor.argtypes = [ctypes.c_int, P, P, P, P, P]
        lib.kcfd_oracle_ilu0_factor.restype = ctypes.c_int
        lib.kcfd_oracle_ilu0_solve.argtypes = [ctypes.c_int, P, P, P, P, P, P]
        lib.kcfd_oracle_ilu0_solve.restype = None
        lib.kcfd_oracle_color_order.argtypes = [ctypes.c_int, P, P, P, P]
        lib.kcfd_oracle_color_order.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class ILU0:
    """Structural ILU(0) factors on A's canonical CSR pattern (C oracle)."""

    def __init__(self, A):
        A = sparse.csr_matrix(A)
        A.sort_indices()
        self.n = A.shape[0]
        self.indptr = np.ascontiguousarray(A.indptr, np.int32)
        self.indices = np.ascontiguousarray(A.indices, np.int32)
        a = np.ascontiguousarray(A.data, np.float64)
        self.lu = np.empty_like(a)
        self.diagpos = np.full(self.n, -1, np.int32)
        st = _lib().kcfd_oracle_ilu0_factor(self.n, _p(self.indptr), _p(self.indices),
                                            _p(a), _p(self.lu), _p(self.diagpos))
        if st != 0:
            raise ValueError(f"ILU(0) breakdown at row {-st - 1}")

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        _lib().kcfd_oracle_ilu0_solve(self.n, _p(self.indptr), _p(self.indices), _p(self.lu),
                                      _p(self.diagpos), _p(b), _p(x))
        return x

    def factors(self):
        """(L, U) as scipy CSR, L unit lower."""
        n = self.n
        rows = np.repeat(np.arange(n), np.diff(self.indptr))
        cols = self.indices
        low = cols < rows
        L = sparse.csr_matrix((self.lu[low], (rows[low], cols[low])), shape=(n, n)) + sparse.eye(n)
        up = cols >= rows
        U = sparse.csr_matrix((self.lu[up], (rows[up], cols[up])), shape=(n, n))
        return L.tocsr(), U.tocsr()


def ilu0_python(A):
    """Pure-Python ILU(0), IKJ (Saad Alg. 10.4), for tiny matrices; used to
    cross-check the C oracle bit for bit."""
    A = sparse.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    ip, ix = A.indptr, A.indices
    lu = A.data.astype(np.float64).copy()
    dpos = np.full(n, -1, np.int64)
    for i in range(n):
        pos = {int(ix[q]): q for q in range(ip[i], ip[i + 1])}
        for q in range(ip[i], ip[i + 1]):
            k = int(ix[q])
            if k >= i:
                break
            lik = lu[q] / lu[dpos[k]]
            lu[q] = lik
            for r in range(dpos[k] + 1, ip[k + 1]):
                j = int(ix[r])
                if j in pos:
                    lu[pos[j]] -= lik * lu[r]
        dpos[i] = pos[i]
    return lu, dpos


def color_order(A):
    """First-fit colouring of A + A^T (ilu0.c): (perm sorted by colour,
    colours per row, number of colours)."""
    A = sparse.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    ip = np.ascontiguousarray(A.indptr, np.int32)
    ix = np.ascontiguousarray(A.indices, np.int32)
    perm = np.empty(n, np.int32)
    col = np.empty(n, np.int32)
    nc = _lib().kcfd_oracle_color_order(n, _p(ip), _p(ix), _p(perm), _p(col))
    if nc < 0:
        raise MemoryError("colour ordering")
    return perm, col, nc


def permuted(A, perm):
    """P A P^T as canonical CSR (row k = row perm[k] of A)."""
    A = sparse.csr_matrix(A)
    Ap = sparse.csr_matrix(A[perm][:, perm])
    Ap.sort_indices()
    return Ap


class ILU0Ordered:
    """M^{-1} z = P^T ILU0(P A P^T)^{-1} P z for a row ordering perm."""

    def __init__(self, A, perm):
        self.perm = np.asarray(perm, np.int64)
        self.inv = np.empty_like(self.perm)
        self.inv[self.perm] = np.arange(len(self.perm))
        self.f = ILU0(permuted(A, self.perm))

    def solve(self, z):
        return self.f.solve(np.asarray(z, np.float64)[self.perm])[self.inv]


def make_preconditioner(A, method: str):
    """core/preconditioner.py:40-54 (+ 'ilu0_struct', 'ilu0_color')."""
    n = A.shape[0]
    if method == "none":
        return None
    if method == "jacobi":
        d = A.diagonal()
        with np.errstate(divide="ignore"):
            dinv = np.where(np.abs(d) > 1e-30, 1.0 / d, 0.0)
        return LinearOperator((n, n), matvec=lambda x: dinv * x)
    if method == "ilu0":
        ilu = spilu(sparse.csc_matrix(A), fill_factor=1, drop_tol=0)
        return LinearOperator((n, n), matvec=ilu.solve)
    if method == "ilu0_struct":
        f = ILU0(A)
        return LinearOperator((n, n), matvec=f.solve)
    if method == "ilu0_color":
        f = ILU0Ordered(A, color_order(A)[0])
        return LinearOperator((n, n), matvec=f.solve)
    raise ValueError(f"Unknown preconditioner: {method}")


def solve_bicgstab(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="jacobi", M=None):
    """scipy BiCGSTAB with the reference's call; returns (x, info, iterations)
    where iterations = number of callback calls = completed full iterations."""
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    if x0 is None:
        x0 = np.zeros(n)
    if M is None:
        M = make_preconditioner(A, precond)
    count = [0]

    def cb(_xk):
        count[0] += 1

    x, info = bicgstab(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol,
                       maxiter=maxiter, M=M, callback=cb)
    return x, info, count[0]


def solve_cg(A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=None, precond="jacobi", M=None):
    """scipy CG with the call of solve_pressure_correction
    (core/linear_solver.py:88-93); returns (x, info, iterations), iterations
    = callback calls = completed iterations."""
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    if x0 is None:
        x0 = np.zeros(n)
    if M is None:
        M = make_preconditioner(A, precond)
    count = [0]

    def cb(_xk):
        count[0] += 1

    x, info = cg(A, b, x0=np.array(x0, np.float64), rtol=rtol, atol=atol, maxiter=maxiter, M=M,
                 callback=cb)
    return x, info, count[0]


def true_rel_residual(A, x, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(b - A @ x) / (nb if nb > 0 else 1.0))
