"""The CG oracle (scipy cg with the call of solve_pressure_correction,
core/linear_solver.py:88-93) pinned by known answers: CG ends in at most k
iterations on an SPD matrix with k distinct eigenvalues, Jacobi-preconditioned
CG is exact in one iteration on a diagonal matrix, and the solution of an SPD
pressure-correction system equals the direct solve."""

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import spsolve

from helpers import oracle_system
from oracle import solve as osolve


def test_cg_distinct_eigenvalues():
    rng = np.random.default_rng(3)
    d = np.repeat([1.0, 3.0, 10.0, 40.0], 25)
    Q, _ = np.linalg.qr(rng.normal(size=(100, 100)))
    A = sparse.csr_matrix(Q @ np.diag(d) @ Q.T)
    b = rng.uniform(-1, 1, 100)
    x, info, it = osolve.solve_cg(A, b, rtol=1e-12, precond="none")
    assert info == 0 and it <= 4
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-11


def test_cg_jacobi_exact_on_diagonal():
    rng = np.random.default_rng(4)
    A = sparse.diags(rng.uniform(1, 100, 300)).tocsr()
    b = rng.uniform(-1, 1, 300)
    x, info, it = osolve.solve_cg(A, b, rtol=1e-12, precond="jacobi")
    assert info == 0 and it == 1
    assert np.allclose(x, b / A.diagonal(), rtol=1e-15, atol=0)


def test_cg_pressure_matches_direct():
    from kcfd_amd import workloads
    A, b, _ = oracle_system(workloads.hex_pressure(8, seed=1002))
    assert abs(A - A.T).max() == 0.0               # the SIMPLE pressure matrix is symmetric
    x, info, it = osolve.solve_cg(A, b, rtol=1e-12, precond="jacobi")
    xd = spsolve(A.tocsc(), b)
    assert info == 0 and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_cg_maxiter_reports_maxiter():
    """scipy's loop exhausted: info = maxiter (the GPU CG mirrors it)."""
    from kcfd_amd import workloads
    A, b, _ = oracle_system(workloads.hex_pressure(8, seed=1002))
    x, info, it = osolve.solve_cg(A, b, rtol=1e-12, maxiter=5, precond="none")
    assert info == 5 and it == 5
