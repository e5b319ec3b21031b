"""Pin the textbook ILU(0) oracle (C) used for the GPU ILU(0) parity tests."""

import numpy as np
import pytest
from scipy import sparse
from scipy.sparse.linalg import splu

from oracle import solve as osolve


def _grid_matrix(n, seed=0):
    from kcfd_amd import workloads as W
    from helpers import oracle_system
    A, b, _ = oracle_system(W.hex_pressure(n, seed=seed))
    return A, b


def test_c_and_python_agree_bitwise():
    A, _ = _grid_matrix(4)
    f = osolve.ILU0(A)
    lu, dpos = osolve.ilu0_python(A)
    assert np.array_equal(f.lu, lu)


def test_lu_reproduces_a_on_pattern():
    A, _ = _grid_matrix(5)
    f = osolve.ILU0(A)
    Lm, Um = f.factors()
    P = (Lm @ Um).tocsr()
    A = A.tocsr()
    A.sort_indices()
    # (LU)_ij == A_ij on pattern(A)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    vals = np.asarray(P[rows, A.indices]).ravel()
    assert np.allclose(vals, A.data, rtol=1e-12, atol=1e-14)


def test_tridiagonal_is_exact_lu():
    """Pattern closed under elimination: ILU(0) = LU (natural order, no pivoting)."""
    n = 50
    rng = np.random.default_rng(1)
    main = 4 + rng.uniform(0, 1, n)
    A = sparse.diags([-1 - rng.uniform(0, 0.5, n - 1), main, -1 - rng.uniform(0, 0.5, n - 1)],
                     [-1, 0, 1]).tocsr()
    f = osolve.ILU0(A)
    Lm, Um = f.factors()
    ex = splu(A.tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0,
              options={"SymmetricMode": True})
    assert np.allclose(Lm.toarray(), ex.L.toarray(), rtol=1e-13)
    assert np.allclose(Um.toarray(), ex.U.toarray(), rtol=1e-13)
    b = rng.normal(0, 1, n)
    x, info, it = osolve.solve_bicgstab(A, b, rtol=1e-12, precond="ilu0_struct")
    assert info == 0 and it <= 1


def test_missing_pivot_raises():
    A = sparse.csr_matrix(np.array([[0.0, 1.0], [1.0, 1.0]]))
    with pytest.raises(ValueError):
        osolve.ILU0(A)


def test_ilu_iterations_below_jacobi():
    A, b = _grid_matrix(8, seed=3)
    _, i1, it_j = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="jacobi")
    _, i2, it_i = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="ilu0_struct")
    assert i1 == 0 and i2 == 0 and it_i < it_j
