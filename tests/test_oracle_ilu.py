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


# ----------------------------------------------------- colour ordering (P A P^T)

def _parity(shape):
    idx = np.indices(shape[::-1]).reshape(len(shape), -1)[::-1]   # x fastest
    return idx.sum(0) % 2


def test_color_order_is_grid_parity():
    """First-fit in natural order on 5- and 7-point grids = checkerboard."""
    from kcfd_amd import workloads as W
    from helpers import oracle_system
    A, _ = _grid_matrix(6)
    perm, col, nc = osolve.color_order(A)
    assert nc == 2 and np.array_equal(col, _parity((6, 6, 6)))
    A2, _, _ = oracle_system(W.quad_convdiff(12, seed=1001))
    perm, col, nc = osolve.color_order(A2)
    assert nc == 2 and np.array_equal(col, _parity((12, 12)))
    # perm: rows sorted by colour, ascending inside a colour
    assert np.array_equal(perm, np.lexsort((np.arange(len(col)), col)))


def test_color_order_first_fit_on_tets():
    """Valid colouring (no coupled pair shares a colour) and first-fit: every
    smaller colour is held by an earlier neighbour."""
    from kcfd_amd import workloads as W
    from helpers import oracle_system
    A, _, _ = oracle_system(W.tet_twofluid(6, seed=1003))
    perm, col, nc = osolve.color_order(A)
    G = (abs(A) + abs(A.T)).tocsr()
    G.setdiag(0)
    G.eliminate_zeros()
    r = np.repeat(np.arange(G.shape[0]), np.diff(G.indptr))
    assert not np.any(col[r] == col[G.indices])
    assert nc >= 2
    for i in range(G.shape[0]):
        nb = G.indices[G.indptr[i]:G.indptr[i + 1]]
        earlier = set(col[nb[nb < i]].tolist())
        assert set(range(col[i])) <= earlier


def test_red_black_ilu0_closed_form():
    """Two colours: P A P^T = [[Dr, B], [C, Db]] with Dr, Db diagonal, and
    ILU(0) is L = [[I, 0], [C Dr^-1, I]], U = [[Dr, B], [0, diag(Db - C Dr^-1 B)]]."""
    A, _ = _grid_matrix(5)
    perm, col, nc = osolve.color_order(A)
    assert nc == 2
    Ap = osolve.permuted(A, perm)
    f = osolve.ILU0(Ap)
    Lm, Um = f.factors()
    nr = int(np.sum(col == 0))
    Ad = Ap.toarray()
    Dr, B, C, Db = Ad[:nr, :nr], Ad[:nr, nr:], Ad[nr:, :nr], Ad[nr:, nr:]
    assert np.count_nonzero(Dr - np.diag(np.diag(Dr))) == 0
    assert np.count_nonzero(Db - np.diag(np.diag(Db))) == 0
    CDi = C / np.diag(Dr)[None, :]
    Lx = np.eye(Ad.shape[0])
    Lx[nr:, :nr] = CDi
    Ux = np.zeros_like(Ad)
    Ux[:nr, :nr] = Dr
    Ux[:nr, nr:] = B
    Ux[nr:, nr:] = np.diag(np.diag(Db) - np.einsum("ij,ji->i", CDi, B))
    assert np.allclose(Lm.toarray(), Lx, rtol=1e-14, atol=0)
    assert np.allclose(Um.toarray(), Ux, rtol=1e-13, atol=1e-300)


def test_color_ordered_preconditioner_converges():
    A, b = _grid_matrix(10)
    x, info, it = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="ilu0_color")
    assert info == 0 and osolve.true_rel_residual(A, x, b) < 1e-8
    _, _, it_nat = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="ilu0_struct")
    _, _, it_jac = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="jacobi")
    assert it_nat <= it <= it_jac
