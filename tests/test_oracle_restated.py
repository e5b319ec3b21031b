"""Pin the restated BiCGSTAB and CG (oracle/solve.py) to scipy itself.

With numpy's dot product the restatements must be scipy bit for bit: the
same x (np.array_equal), info and callback count, on systems that reach
every exit of scipy's loops -- convergence, the half-step exit, the loop
exhausted at maxiter, a rho breakdown (case 16 at the reference figure's
settings), an omega breakdown (a zero Jacobi diagonal), a zero right-hand
side and a nonzero x0 -- with none / Jacobi / ILU(0) preconditioners in the
natural and the colour ordering.  The GPU tests then hold the device's FP64
solves to these restatements evaluated in the device's dot-product order
(helpers.dot_tile_order), which a literal loop over the device's threads
pins below.
"""

import math

import numpy as np
import pytest
from scipy import sparse

from helpers import dot_tile_order, oracle_system
from oracle import golden, solve as osolve


def _kc():
    from kcfd_amd import workloads
    return workloads


def _systems():
    w = _kc()
    out = {}
    for n in (16, 32):
        out[f"quad{n}"] = oracle_system(w.quad_convdiff(n, seed=1001))[:2]
    out["hex8"] = oracle_system(w.hex_pressure(8, seed=1002))[:2]
    out["hex10mom"] = oracle_system(w.hex_momentum(10, seed=1002))[:2]
    out["tet"] = oracle_system(w.tet_twofluid(600, seed=1003))[:2]
    return out


SYSTEMS = None


def systems():
    global SYSTEMS
    if SYSTEMS is None:
        SYSTEMS = _systems()
    return SYSTEMS


def _scipy_bicgstab(A, b, **kw):
    x, info, it = osolve.solve_bicgstab(A, b, **kw)
    return x, info, it


@pytest.mark.parametrize("name", ["quad16", "quad32", "hex8", "hex10mom", "tet"])
@pytest.mark.parametrize("pc", ["none", "jacobi", "ilu0_struct", "ilu0_color"])
def test_bicgstab_restated_is_scipy(name, pc):
    A, b = systems()[name]
    xs, infos, its = _scipy_bicgstab(A, b, rtol=1e-8, precond=pc)
    xr, infor, itr, half = osolve.bicgstab_restated(A, b, rtol=1e-8, precond=pc)
    assert (infor, itr) == (infos, its)
    assert np.array_equal(xr, xs)


def test_bicgstab_restated_exits():
    # the reference figure's settings: the Jacobi run ends in a rho breakdown
    A, b = golden.case16_system()
    for pc in ("none", "jacobi"):
        xs, infos, its = _scipy_bicgstab(A, b, rtol=0.0, atol=1e-10, maxiter=2000, precond=pc)
        xr, infor, itr, _ = osolve.bicgstab_restated(A, b, rtol=0.0, atol=1e-10, maxiter=2000, precond=pc)
        assert (infor, itr) == (infos, its) and np.array_equal(xr, xs)
        if pc == "jacobi":
            assert infor == -10
    # loop exhausted
    A, b = systems()["quad16"]
    xs, infos, its = _scipy_bicgstab(A, b, rtol=1e-14, maxiter=3, precond="jacobi")
    xr, infor, itr, _ = osolve.bicgstab_restated(A, b, rtol=1e-14, maxiter=3, precond="jacobi")
    assert infor == infos == 3 and itr == its == 3 and np.array_equal(xr, xs)
    # nonzero x0
    x0 = np.random.default_rng(5).uniform(-1, 1, len(b))
    xs, infos, its = _scipy_bicgstab(A, b, x0=x0, rtol=1e-8, precond="jacobi")
    xr, infor, itr, _ = osolve.bicgstab_restated(A, b, x0=x0, rtol=1e-8, precond="jacobi")
    assert (infor, itr) == (infos, its) and np.array_equal(xr, xs)
    # half-step exit: 4 x = 2 stops at the test on ||s|| before any callback
    one = sparse.csr_matrix(np.array([[4.0]]))
    xr, infor, itr, half = osolve.bicgstab_restated(one, np.array([2.0]), rtol=1e-12)
    xs, infos, its = _scipy_bicgstab(one, np.array([2.0]), rtol=1e-12, precond="none")
    assert (infor, itr, half) == (0, 0, 1) and (infos, its) == (0, 0) and np.array_equal(xr, xs)
    # omega breakdown: a zero on the diagonal gives Jacobi 0 there
    Zd = sparse.csr_matrix((np.array([0.0, 1.0, 1.0, 1.0]), np.array([0, 1, 0, 1]), np.array([0, 2, 4])),
                           shape=(2, 2))
    xs, infos, its = _scipy_bicgstab(Zd, np.array([1.0, 2.0]), rtol=1e-12, maxiter=50, precond="jacobi")
    xr, infor, itr, _ = osolve.bicgstab_restated(Zd, np.array([1.0, 2.0]), rtol=1e-12, maxiter=50,
                                                 precond="jacobi")
    assert infor == infos == -11 and itr == its and np.array_equal(xr, xs)
    # zero right-hand side
    xr, infor, itr, _ = osolve.bicgstab_restated(A, np.zeros(len(b)), rtol=1e-8, precond="jacobi")
    assert (infor, itr) == (0, 0) and not xr.any()


@pytest.mark.parametrize("name", ["quad16", "hex8", "tet"])
@pytest.mark.parametrize("pc", ["none", "jacobi", "ilu0_struct", "ilu0_color"])
def test_cg_restated_is_scipy(name, pc):
    A, b = systems()[name]
    if name == "quad16":      # CG needs the symmetric pressure systems; the
        A = (A + A.T).tocsr()  # symmetric part of the quad operator stands in
        A.sort_indices()
    xs, infos, its = osolve.solve_cg(A, b, rtol=1e-8, precond=pc)
    xr, infor, itr = osolve.cg_restated(A, b, rtol=1e-8, precond=pc)
    assert (infor, itr) == (infos, its)
    assert np.array_equal(xr, xs)
    xs, infos, its = osolve.solve_cg(A, b, rtol=1e-14, maxiter=4, precond=pc)
    xr, infor, itr = osolve.cg_restated(A, b, rtol=1e-14, maxiter=4, precond=pc)
    assert infor == infos == 4 and itr == its == 4 and np.array_equal(xr, xs)


def _dot_literal(a, b, R, tile_key=None):
    """The device's reduction written as loops over tiles, threads, waves
    and shuffle steps; ``tile_key`` (one value per tile) sorts the tile
    list, as a permuted system's tiles are listed by the row of A they
    start at."""
    n = len(a)
    T = 256 * R
    parts = []
    starts = list(range(0, max(n, 1), T))
    if tile_key is not None:
        starts = [t0 for _, t0 in sorted(zip(tile_key, starts))]
    for t0 in starts:
        lane = [0.0] * 256
        for t in range(256):
            acc = 0.0
            for r in range(R):
                i = t0 + r * 256 + t
                if i < n:
                    acc = acc + float(a[i]) * float(b[i])
            lane[t] = acc
        parts.append(_wg_sum(lane))
    fin = [0.0] * 1024
    for j in range(1024):
        acc = 0.0
        for q in range(j, len(parts), 1024):
            acc = acc + parts[q]
        fin[j] = acc
    return _wg_sum(fin)


def _wg_sum(vals):
    waves = []
    for w in range(len(vals) // 64):
        v = list(vals[64 * w:64 * w + 64])
        for o in (32, 16, 8, 4, 2, 1):
            v = [v[l] + v[l ^ o] for l in range(64)]
        waves.append(v[0])
    s = waves[0]
    for q in waves[1:]:
        s = s + q
    return s


@pytest.mark.parametrize("R", [1, 2, 4])
@pytest.mark.parametrize("n", [1, 300, 5000])
def test_dot_tile_order_literal(R, n):
    rng = np.random.default_rng(n + R)
    a = rng.normal(size=n) * np.exp(rng.uniform(-20, 20, n))
    b = rng.normal(size=n)
    d = dot_tile_order(R)(a, b)
    assert d == _dot_literal(a, b, R)
    assert abs(d - math.fsum((a * b).tolist())) <= 1e-12 * np.sum(np.abs(a * b))
    perm = rng.permutation(n)
    key = [int(perm[t0]) for t0 in range(0, n, 256 * R)]
    assert dot_tile_order(R, perm)(a, b) == _dot_literal(a[perm], b[perm], R, key)
    assert dot_tile_order(R, tiles_by=perm)(a[perm], b[perm]) == _dot_literal(a[perm], b[perm], R, key)


def test_dot_tile_order_is_an_order():
    """Integers sum exactly in any order; a cancelling pair in different
    tiles survives only in the order that adds it first."""
    v = np.arange(10000, dtype=np.float64)
    assert dot_tile_order(2)(v, np.ones_like(v)) == float(v.sum())
    a = np.zeros(2048)
    a[0], a[1], a[1024] = 1e16, 1.0, -1e16
    # tile 0 (rows 0..511) holds 1e16 + 1 = 1e16 (rounded); tile 2 holds -1e16
    assert dot_tile_order(2)(a, np.ones_like(a)) == 0.0
    assert math.fsum(a.tolist()) == 1.0


@pytest.mark.parametrize("solver", ["bicgstab", "cg"])
@pytest.mark.parametrize("pc", ["jacobi", "ilu0_color"])
def test_restated_in_row_order_equals_gathered(solver, pc):
    """helpers.restated_device runs a permuted solve in P-order (as the
    device does); it equals the restatement in A's order with every dot
    product's terms summed in P-order, bit for bit."""
    from helpers import restated_device
    A, b = systems()["hex8"] if solver == "cg" else systems()["hex10mom"]
    perm = osolve.color_order(A)[0]
    run = osolve.cg_restated if solver == "cg" else osolve.bicgstab_restated
    for R in (1, 2, 4):
        ref = run(A, b, rtol=1e-8, precond=pc, dot=dot_tile_order(R, perm))
        got = restated_device(A, b, solver=solver, rows_per_thread=R, perm=perm, rtol=1e-8, precond=pc)
        assert np.array_equal(ref[0], got[0]) and ref[1:] == got[1:]
