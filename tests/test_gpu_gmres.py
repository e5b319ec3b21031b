"""GPU GMRES (solve_linear_system(method='gmres'), core/linear_solver.py:
64-68) against the GMRES oracle.  The device's solve equals the restated
GMRES (oracle.solve.gmres_restated, itself equal to scipy's gmres in
scipy's order) evaluated in the device's summation order: same inner
iterations, cycles and info, and x bit for bit.  scipy's own result and the
band the restatement spans under other summation orders of the dot
products and of the x update (the one arithmetic the device does in another
order) are printed beside it; FP64-verified residual below the tolerance.  Known answers
(distinct eigenvalues, exact preconditioner, maxiter = restart cycles, zero
right-hand side, x0 already converged); the drop-in keeps the reference's
call semantics (atol = tol, rtol = 1e-5)."""

import numpy as np
import pytest
from scipy import sparse
from scipy.sparse.linalg import LinearOperator

from helpers import dot_device_order, gmres_band, oracle_system, rel_l2
from oracle import solve as osolve

pytestmark = pytest.mark.gpu


def _case(which):
    from kcfd_amd import workloads
    return {"hex12": lambda: workloads.hex_pressure(12, seed=1002),
            "quad32": lambda: workloads.quad_convdiff(32, seed=1001),
            "tet6": lambda: workloads.tet_twofluid(6, seed=1003)}[which]()


def _device_order_restated(A, b, restart=None, perm=None, **kw):
    """The restated GMRES in the device's summation order.  With a row
    ordering perm (colour-ordered ILU(0)) it runs, as the device does, on
    P A P^T with each row's entries in A's CSR order, b and x0 gathered, the
    ILU(0) of P A P^T applied in that order, and x scattered back."""
    from kcfd_amd import _lib
    dot = dot_device_order(_lib.default_context().info()["cus"])
    rk = {k: v for k, v in kw.items() if k in ("x0", "rtol", "atol", "maxiter", "precond")}
    if perm is None:
        return osolve.gmres_restated(A, b, restart=restart or 20, dot=dot, xupdate="sequential", **rk)
    A = sparse.csr_matrix(A)
    n = A.shape[0]
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    lens = np.diff(A.indptr)[perm]
    ip = np.concatenate([[0], np.cumsum(lens)])
    rows = [np.arange(A.indptr[r], A.indptr[r + 1]) for r in perm]
    q = np.concatenate(rows)
    Ap = sparse.csr_matrix((A.data[q], inv[A.indices[q]], ip), shape=A.shape)   # entry order kept
    f = osolve.ILU0Ordered(A, perm).f
    M = LinearOperator((n, n), matvec=f.solve)
    rk.pop("precond", None)
    if "x0" in rk:
        rk["x0"] = np.asarray(rk["x0"])[perm]
    xp, info, it, cyc = osolve.gmres_restated(Ap, np.asarray(b)[perm], restart=restart or 20, dot=dot,
                                              xupdate="sequential", M=M, **rk)
    return xp[inv], info, it, cyc


def _check(x, res, A, b, label, restart=None, perm=None, **kw):
    """The bar: the restated GMRES summed in the device's own order gives
    the same iterations, cycles and info and the same x, bit for bit.  The
    band of scipy and the restatement under four other summation orders is
    printed beside it (long restarted runs move by a few iterations between
    orders: hex12 unpreconditioned takes 1283 in scipy's order and 1280 in
    the device's)."""
    its, cycs, infos, xo, spread = gmres_band(A, b, restart=restart, **kw)
    xd, info_d, it_d, cyc_d = _device_order_restated(A, b, restart=restart, perm=perm, **kw)
    its, cycs, infos = its + [it_d], cycs + [cyc_d], infos + [info_d]
    spread = max(spread, rel_l2(xd, xo))
    d = rel_l2(x, xo)
    dd = rel_l2(x, xd)
    print(f"{label}: gpu iters={res.iters} cycles={res.refinements} info={res.info} "
          f"resid={res.rel_resid:.2e} | oracle {its[0]} it / {cycs[0]} cycles (summation-order band "
          f"{min(its)}..{max(its)} it, {min(cycs)}..{max(cycs)} cycles) | relL2={d:.2e} "
          f"(spread {spread:.2e}) | device-order restatement {it_d} it / {cyc_d} cycles, relL2={dd:.2e}")
    assert (res.iters, res.refinements, res.info) == (it_d, cyc_d, info_d)
    assert np.array_equal(x, xd), dd
    return d


@pytest.mark.parametrize("which", ["hex12", "quad32", "tet6"])
@pytest.mark.parametrize("pc", ["none", "jacobi", "ilu0"])
def test_gmres_matches_oracle(ctx, which, pc):
    from kcfd_amd import device
    A, b, _ = oracle_system(_case(which))
    M = device.DeviceMatrix.from_csr(ctx, A)
    x, res = M.solve(b, rtol=1e-9, precond=pc, method="gmres")
    assert res.info == 0 and res.rel_resid < 1e-9
    assert osolve.true_rel_residual(A, x, b) < 1e-9
    _check(x, res, A, b, f"gmres {which} {pc}", rtol=1e-9,
           precond="ilu0_struct" if pc == "ilu0" else pc)


@pytest.mark.parametrize("restart", [5, 30, 64])
def test_gmres_restart_lengths(ctx, restart):
    from kcfd_amd import device
    A, b, _ = oracle_system(_case("quad32"))
    x, res = device.DeviceMatrix.from_csr(ctx, A).solve(b, rtol=1e-9, precond="jacobi",
                                                        method="gmres", restart=restart)
    _check(x, res, A, b, f"gmres quad32 jacobi m={restart}", restart=restart, rtol=1e-9,
           precond="jacobi")


def test_gmres_colour_ilu0(ctx):
    """ILU(0) in the colour ordering: the solve runs in P-order (b, x0
    gathered, x scattered)."""
    from kcfd_amd import device
    A, b, _ = oracle_system(_case("hex12"))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="color")
    x0 = np.random.default_rng(3).uniform(-1, 1, A.shape[0])
    x, res = M.solve(b, x0=x0, rtol=1e-9, precond="ilu0", method="gmres")
    assert res.info == 0
    perm, _ = M.ilu0_order()
    _check(x, res, A, b, "gmres hex12 ilu0 colour, x0 != 0", rtol=1e-9, precond="ilu0_color", x0=x0,
           perm=np.asarray(perm, np.int64))


def test_gmres_known_answers(ctx):
    """Four distinct eigenvalues: at most 4 inner iterations; Jacobi on a diagonal matrix: exact after 1;
    maxiter = 2 cycles of 10: info 2 after 20 inner iterations, x equal to
    scipy's; b = 0: x = 0, no iterations; x0 already converged: no
    iterations."""
    from kcfd_amd import device
    rng = np.random.default_rng(5)
    A = sparse.diags(np.repeat([1.0, 3.0, 10.0, 25.0], 100)).tocsr()
    b = rng.uniform(-1, 1, 400)
    x, res = device.DeviceMatrix.from_csr(ctx, A).solve(b, rtol=1e-12, precond="none", method="gmres")
    assert res.info == 0 and res.iters <= 4 and rel_l2(A @ x, b) < 1e-11
    D = sparse.diags(rng.uniform(1, 100, 1000)).tocsr()
    bd = rng.uniform(-1, 1, 1000)
    x, res = device.DeviceMatrix.from_csr(ctx, D).solve(bd, rtol=1e-12, precond="jacobi", method="gmres")
    assert res.info == 0 and res.iters == 1
    assert rel_l2(x, bd / D.diagonal()) < 1e-15
    A2, b2, _ = oracle_system(_case("quad32"))
    M2 = device.DeviceMatrix.from_csr(ctx, A2)
    x, res = M2.solve(b2, rtol=1e-12, maxiter=2, precond="none", method="gmres", restart=10)
    xo, info_o, it_o = osolve.solve_gmres(A2, b2, rtol=1e-12, maxiter=2, restart=10)
    assert res.info == info_o == 2 and res.iters == it_o == 20 and res.refinements == 2
    assert rel_l2(x, xo) < 1e-10
    x, res = M2.solve(np.zeros(A2.shape[0]), rtol=1e-8, precond="none", method="gmres")
    assert res.info == 0 and res.iters == 0 and not x.any()
    xs = osolve.solve_gmres(A2, b2, rtol=1e-12)[0]
    x, res = M2.solve(b2, x0=xs, rtol=1e-8, precond="none", method="gmres")
    assert res.info == 0 and res.iters == 0 and np.array_equal(x, xs)


def test_gmres_dropin(ctx, capsys):
    """solve_linear_system(method='gmres') with the reference's call:
    atol = tol, scipy's default rtol = 1e-5, maxiter = restart cycles; a
    non-converged solve prints the reference's message and returns x."""
    from kcfd_amd.core.fvm_operators import FVMSystem
    from kcfd_amd.core.linear_solver import solve_linear_system
    A, b, _ = oracle_system(_case("quad32"))
    sysm = FVMSystem(A.shape[0])
    coo = A.tocoo()
    for i, j, v in zip(coo.row.tolist(), coo.col.tolist(), coo.data.tolist()):
        if i == j:
            sysm.add_diagonal(i, v)
        else:
            sysm.add_off_diagonal(i, j, v)
    sysm.rhs = b.copy()
    for pc in ("none", "jacobi", "ilu0"):
        x, info = solve_linear_system(sysm, method="gmres", tol=1e-8, preconditioner=pc, return_info=True)
        opc = "ilu0_struct" if pc == "ilu0" else pc
        xo, info_o, it_o = osolve.solve_gmres(A, b, rtol=1e-5, atol=1e-8, maxiter=1000, precond=opc)
        print(f"drop-in gmres {pc}: gpu {info['iters']} it, oracle {it_o}; relL2 {rel_l2(x, xo):.2e}")
        assert info["info"] == info_o == 0 and abs(info["iters"] - it_o) <= 2
        assert np.linalg.norm(b - A @ x) <= max(1e-8, 1e-5 * np.linalg.norm(b))
    x = solve_linear_system(sysm, method="gmres", tol=1e-14, maxiter=1, preconditioner="none",
                            rtol=0.0)
    assert "[gmres]" in capsys.readouterr().out and x.shape == b.shape


def test_gmres_graph_replay_equals_direct(ctx, monkeypatch):
    """From the second cycle on the solve replays a captured hipGraph of one
    cycle; KCFD_GMRES_GRAPH=0 launches every cycle directly.  Both give the
    same iterates bit for bit over every replay (and twice in a row with the
    graph), in the natural and the colour-ordered ILU(0)."""
    from kcfd_amd import device
    A, b, _ = oracle_system(_case("quad32"))
    for pc, order in (("jacobi", None), ("ilu0", "color")):
        M = device.DeviceMatrix.from_csr(ctx, A)
        M.setup_precond(pc, "fp64", order=order)
        runs = []
        for mode in ("1", "0", "1"):
            monkeypatch.setenv("KCFD_GMRES_GRAPH", mode)
            x, res = M.solve(b, rtol=1e-10, precond=pc, method="gmres", restart=8)
            runs.append((x, res.iters, res.refinements))
        print(f"graph vs direct {pc}: {runs[0][1]} iterations in {runs[0][2]} cycles")
        assert runs[0][2] > 3
        for x, it, cyc in runs[1:]:
            assert (it, cyc) == runs[0][1:] and np.array_equal(x, runs[0][0])


def test_invalid_options_raise(ctx):
    """The C-ABI rejects what it cannot run, with a message: GMRES in mixed
    precision, a restart above 64, GMRES in a batch, and rows per thread
    other than 1, 2 or 4; the matrix stays usable afterwards."""
    from kcfd_amd import device
    from kcfd_amd._lib import DevBuf, KcfdError
    A, b, _ = oracle_system(_case("quad32"))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("jacobi", "mixed")
    with pytest.raises(KcfdError, match="FP64"):
        M.solve(b, rtol=1e-8, precond="jacobi", precision="mixed", method="gmres")
    with pytest.raises(KcfdError, match="restart"):
        M.solve(b, rtol=1e-8, precond="jacobi", method="gmres", restart=65)
    with pytest.raises(KcfdError, match="rows_per_thread"):
        M.solve(b, rtol=1e-8, precond="jacobi", rows_per_thread=3)
    bd, xd = DevBuf.from_array(ctx, b), DevBuf.from_array(ctx, np.zeros(len(b)))
    with pytest.raises(KcfdError, match="GMRES"):
        device.Batch(ctx, [M], [bd], [xd], device.default_opts(rtol=1e-8, precond="jacobi", method="gmres"))
    x, res = M.solve(b, rtol=1e-8, precond="jacobi", method="gmres")
    assert res.info == 0
