"""The GMRES oracle (scipy gmres with the call of solve_linear_system,
core/linear_solver.py:64-68) and its step-by-step restatement, pinned by
known answers: GMRES without restarts is exact after k inner iterations on a
diagonalisable matrix with k distinct eigenvalues, after one iteration when
the preconditioner inverts A, the restated algorithm (the device's steps)
reproduces scipy's iteration and cycle counts and solution, ``lartg``
equals LAPACK's dlartg bit for bit (scaled branches included), and maxiter
counts restart cycles."""

import numpy as np
import pytest
from scipy import sparse
from scipy.linalg.lapack import dlartg

from helpers import oracle_system
from oracle import solve as osolve


def test_gmres_distinct_eigenvalues():
    """Nonsymmetric A = X D X^{-1} with 5 distinct eigenvalues: the Krylov
    space closes after 5 steps, inside one cycle of 20."""
    rng = np.random.default_rng(5)
    d = np.repeat([1.0, 2.0, 5.0, 9.0, 20.0], 16)
    X = np.eye(80) + 0.1 * rng.normal(size=(80, 80))
    A = sparse.csr_matrix(X @ np.diag(d) @ np.linalg.inv(X))
    b = rng.uniform(-1, 1, 80)
    x, info, it, pr, cyc = osolve.solve_gmres(A, b, rtol=1e-10, counts=True)
    assert info == 0 and it <= 6 and cyc == 1
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-9
    assert np.all(np.diff(pr) <= 1e-12)        # the minimal residual never grows in a cycle


def test_gmres_jacobi_exact_on_diagonal():
    rng = np.random.default_rng(6)
    A = sparse.diags(rng.uniform(1, 100, 200)).tocsr()
    b = rng.uniform(-1, 1, 200)
    x, info, it = osolve.solve_gmres(A, b, rtol=1e-12, precond="jacobi")
    assert info == 0 and it == 1
    assert np.allclose(x, b / A.diagonal(), rtol=1e-14, atol=0)


@pytest.mark.parametrize("case,pc,restart", [("quad", "none", 20), ("quad", "jacobi", 7),
                                              ("hex", "ilu0_struct", 20), ("hex", "jacobi", 5),
                                              ("tet", "none", 30)])
def test_restated_equals_scipy(case, pc, restart):
    from kcfd_amd import workloads
    c = {"quad": lambda: workloads.quad_convdiff(24, seed=1001),
         "hex": lambda: workloads.hex_pressure(10, seed=1002),
         "tet": lambda: workloads.tet_twofluid(4, seed=1003)}[case]()
    A, b, _ = oracle_system(c)
    xs, info_s, it_s, _, cyc_s = osolve.solve_gmres(A, b, rtol=1e-9, precond=pc, restart=restart,
                                                    counts=True)
    xr, info_r, it_r, cyc_r = osolve.gmres_restated(A, b, rtol=1e-9, precond=pc, restart=restart)
    print(f"{case} {pc} m={restart}: scipy {it_s} it / {cyc_s} cycles, restated {it_r} / {cyc_r}")
    assert info_s == info_r == 0 and it_s == it_r and cyc_s == cyc_r
    assert np.linalg.norm(xr - xs) / np.linalg.norm(xs) < 1e-12


def test_restated_maxiter_counts_cycles():
    from kcfd_amd import workloads
    A, b, _ = oracle_system(workloads.quad_convdiff(32, seed=1001))
    xs, info_s, it_s, _, cyc_s = osolve.solve_gmres(A, b, rtol=1e-12, maxiter=2, restart=10,
                                                    counts=True)
    xr, info_r, it_r, cyc_r = osolve.gmres_restated(A, b, rtol=1e-12, maxiter=2, restart=10)
    assert info_s == info_r == 2 and cyc_s == cyc_r == 2 and it_s == it_r == 20
    assert np.linalg.norm(xr - xs) / np.linalg.norm(xs) < 1e-12


def test_lartg_matches_lapack():
    rng = np.random.default_rng(8)
    vals = [(0.0, 3.0), (0.0, -3.0), (-2.0, 0.0), (3.0, 4.0), (-3.0, 4.0), (1e-300, 1e-301),
            (-1e300, 3e299), (1e-170, -2e-170), (5e160, 1e-5), (1e-200, 1.0)]
    for _ in range(2000):
        f, g = rng.standard_normal(2) * 10.0 ** rng.uniform(-300, 300, 2)
        vals.append((f, g))
    for f, g in vals:
        assert osolve.lartg(f, g) == tuple(dlartg(f, g)), (f, g)


def test_zero_rhs_and_initial_guess():
    from kcfd_amd import workloads
    A, b, _ = oracle_system(workloads.quad_convdiff(12, seed=1001))
    x, info, it, cyc = osolve.gmres_restated(A, np.zeros(len(b)))
    assert info == 0 and it == 0 and not x.any()
    xs = osolve.solve_gmres(A, b, rtol=1e-12)[0]
    x, info, it, cyc = osolve.gmres_restated(A, b, x0=xs, rtol=1e-8)
    assert info == 0 and it == 0                  # x0 already satisfies the tolerance
