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


@pytest.mark.parametrize("which", ["quad12", "hex6", "quad7x5", "quad100"])
def test_segment_order_invariants(which):
    """The segment ordering restated for the device's KCFD_ORDER_SEGMENT,
    against its definition: segments (recomputed here: runs of at most 4
    rows each coupled to the one before) are never coupled to another
    segment of their colour, colour 0 comes first, every segment's rows keep
    their order, and within a colour the rows come in groups of 64 segments
    numbered step-major, longest segments first; on grids in their natural
    numbering two colours suffice."""
    from kcfd_amd import workloads as W
    from helpers import oracle_system
    case = {"quad12": lambda: W.quad_convdiff(12, seed=1001),
            "hex6": lambda: W.hex_pressure(6, seed=1002),
            "quad7x5": lambda: W.pressure_case(W.quad_mesh(7, 5), seed=3),
            "quad100": lambda: W.quad_convdiff(100, seed=1001)}[which]()
    A, _, _ = oracle_system(case)
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    perm, col, nc = osolve.segment_order(A, seg_len=4)
    assert nc == 2
    assert np.array_equal(np.sort(perm), np.arange(n))
    pos = np.empty(n, np.int64)
    pos[perm] = np.arange(n)
    assert np.all(np.diff(col[perm]) >= 0)                      # colour 0 first
    S = (abs(A) + abs(A.T)).tocsr()
    seg = np.zeros(n, np.int64)
    g, length = 0, 0
    for i in range(1, n):
        if length < 3 and S[i, i - 1] != 0:
            length += 1
        else:
            g, length = g + 1, 0
        seg[i] = g
    assert np.all(col == col[np.r_[0, np.flatnonzero(np.diff(seg)) + 1]][seg])   # one colour per segment
    Sc = S.tocoo()
    off = (Sc.row != Sc.col) & (seg[Sc.row] != seg[Sc.col])
    assert not np.any(col[Sc.row[off]] == col[Sc.col[off]])
    same = np.flatnonzero(np.diff(seg) == 0)                    # i, i + 1 in one segment
    assert np.all(pos[same + 1] > pos[same])
    # step-major groups: within a colour, segments 64 at a time; a group
    # lists the first rows of its segments (longest first), then the second
    # rows in the same lane order, ...
    starts = np.r_[0, np.flatnonzero(np.diff(seg)) + 1]
    lens = np.diff(np.r_[starts, n])
    k0 = 0
    for c in range(nc):
        segs = [s_ for s_ in range(len(starts)) if col[starts[s_]] == c]
        for f in range(0, len(segs), 64):
            grp = segs[f:f + 64]
            order = sorted(grp, key=lambda s_: -lens[s_])
            rows = [starts[s_] + k for k in range(max(lens[grp])) for s_ in order if lens[s_] > k]
            assert np.array_equal(perm[k0:k0 + len(rows)], rows)
            k0 += len(rows)
    assert k0 == n


def test_segment_order_cuts_iterations():
    """On a 2-D convection-diffusion grid the segment-ordered ILU(0) keeps
    the couplings along the grid lines and needs fewer BiCGSTAB iterations
    than the row colouring (natural order needs fewer still)."""
    from kcfd_amd import workloads as W
    from helpers import oracle_system
    A, b, _ = oracle_system(W.quad_convdiff(64, seed=6000))
    A = sparse.csr_matrix(A)
    its = {}
    for name, perm in (("colour", osolve.color_order(A)[0]), ("segment", osolve.segment_order(A)[0]),
                       ("natural", np.arange(A.shape[0]))):
        f = osolve.ILU0Ordered(A, perm)
        M = osolve.LinearOperator(A.shape, matvec=f.solve)
        _, info, it = osolve.solve_bicgstab(A, b, rtol=1e-8, M=M)
        assert info == 0
        its[name] = it
    print(its)
    assert its["natural"] < its["segment"] < its["colour"]
