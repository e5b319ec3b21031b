"""GPU CG (solve_pressure_correction, core/linear_solver.py:73-93) against
the oracle: x, info and the iteration count equal, bit for bit, to scipy's
cg restated with the device's dot-product order (oracle/solve.py:
cg_restated, pinned to scipy itself in tests/test_oracle_restated.py), the
FP64-verified residual below the tolerance; known answers; a batch of CG
systems equals each system alone; the drop-in keeps the reference's call
semantics (atol = tol, rtol = 1e-5)."""

import numpy as np
import pytest
from scipy import sparse

from helpers import oracle_system, rel_l2, restated_device

pytestmark = pytest.mark.gpu


def _case(which):
    from kcfd_amd import workloads
    return {"hex12": lambda: workloads.hex_pressure(12, seed=1002),
            "quad32": lambda: workloads.pressure_case(workloads.quad_mesh(32), 1005),
            "tet6": lambda: workloads.pressure_case(workloads.tet_twofluid(6, seed=1003).mesh, 1006)}[which]()


@pytest.mark.parametrize("which", ["hex12", "quad32", "tet6"])
@pytest.mark.parametrize("pc", ["none", "jacobi", "ilu0"])
def test_cg_matches_oracle(ctx, which, pc):
    from kcfd_amd import device
    A, b, _ = oracle_system(_case(which))
    M = device.DeviceMatrix.from_csr(ctx, A)
    x, res = M.solve(b, rtol=1e-8, precond=pc, method="cg")
    opc = "ilu0_struct" if pc == "ilu0" else pc
    _check(x, res, A, b, f"cg {which} {pc}", rtol=1e-8, precond=opc)
    assert res.info == 0 and res.rel_resid < 1e-8


def _check(x, res, A, b, label, perm=None, **kw):
    from oracle import solve as osolve
    xr, info, it = restated_device(A, b, solver="cg", perm=perm, **kw)
    xs, info_s, it_s = osolve.solve_cg(A, b, **kw)
    print(f"{label}: gpu iters={res.iters} info={res.info} resid={res.rel_resid:.2e} | restated "
          f"(device order) {it} info {info} | scipy {it_s}, relL2 to scipy {rel_l2(x, xs):.2e}")
    assert (res.info, res.iters) == (info, it)
    assert np.array_equal(x, xr)


def test_cg_colour_ilu0(ctx):
    """ILU(0) in the colour ordering (the solver then runs in P-order)."""
    from kcfd_amd import device
    A, b, _ = oracle_system(_case("hex12"))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="color")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", method="cg")   # keeps the colour ordering
    assert res.info == 0 and res.rel_resid < 1e-8
    _check(x, res, A, b, "cg colour", perm=M.ilu0_order()[0], rtol=1e-8, precond="ilu0_color")


def test_cg_segment_ilu0(ctx):
    """ILU(0) in the segment ordering (the segment apply kernel, CG's
    variant): equal to the restated CG in that ordering."""
    from kcfd_amd import device
    A, b, _ = oracle_system(_case("hex12"))
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("ilu0", "fp64", order="segment")
    x, res = M.solve(b, rtol=1e-8, precond="ilu0", method="cg")
    assert res.info == 0 and res.rel_resid < 1e-8
    _check(x, res, A, b, "cg segment", perm=M.ilu0_order()[0], rtol=1e-8, precond="ilu0_segment")


def test_cg_known_answers(ctx):
    """Three distinct eigenvalues: at most 3 iterations; Jacobi on a
    diagonal matrix: exact after 1; maxiter reached: info = maxiter."""
    from kcfd_amd import device
    rng = np.random.default_rng(5)
    A = sparse.diags(np.repeat([1.0, 3.0, 10.0], 200)).tocsr()
    b = rng.uniform(-1, 1, 600)
    x, res = device.DeviceMatrix.from_csr(ctx, A).solve(b, rtol=1e-12, precond="none", method="cg")
    assert res.info == 0 and res.iters <= 3 and rel_l2(A @ x, b) < 1e-11
    D = sparse.diags(rng.uniform(1, 100, 1000)).tocsr()
    x, res = device.DeviceMatrix.from_csr(ctx, D).solve(b[:1] * np.ones(1000), rtol=1e-12,
                                                        precond="jacobi", method="cg")
    assert res.info == 0 and res.iters == 1
    A2, b2, _ = oracle_system(_case("hex12"))
    x, res = device.DeviceMatrix.from_csr(ctx, A2).solve(b2, rtol=1e-12, maxiter=5, precond="none",
                                                         method="cg")
    assert res.info == 5 and res.iters == 5


def test_cg_batch_equals_alone(ctx):
    from kcfd_amd import device
    from kcfd_amd._lib import DevBuf
    cases = [_case(w) for w in ("hex12", "quad32", "tet6")]
    mats, bs, xs = [], [], []
    for c in cases:
        A, b, _ = oracle_system(c)
        M = device.DeviceMatrix.from_csr(ctx, A)
        M.setup_precond("ilu0", "fp64", order="color")
        mats.append(M)
        bs.append(DevBuf.from_array(ctx, b))
        xs.append(DevBuf.from_array(ctx, np.zeros(A.shape[0])))
    opts = device.default_opts(rtol=1e-8, precond="ilu0", precision="fp64", method="cg")
    bt = device.Batch(ctx, mats, bs, xs, opts)
    try:
        res = bt.run()
    finally:
        bt.close()
    for M, b, x, r in zip(mats, bs, xs, res):
        xa = DevBuf.from_array(ctx, np.zeros(M.n))
        ra = M.solve_dev(b.p, xa.p, opts, setup=False)
        assert r.info == 0 and ra.iters == r.iters
        assert np.array_equal(xa.download(), x.download())


def test_solve_pressure_correction_dropin(ctx):
    """The reference's call: cg(A, b, x0=0, atol=tol, maxiter, M) with scipy's
    default rtol = 1e-5."""
    from kcfd_amd.core import linear_solver as ls
    from kcfd_amd.core.fvm_operators import FVMSystem
    from oracle import solve as osolve
    A, b, _ = oracle_system(_case("hex12"))
    s = FVMSystem(A.shape[0])
    coo = A.tocoo()
    for i, j, v in zip(coo.row.tolist(), coo.col.tolist(), coo.data.tolist()):
        if i == j:
            s.add_diagonal(i, v)
        else:
            s.add_off_diagonal(i, j, v)
    s.rhs = b.copy()
    x, info = ls.solve_pressure_correction(s, method="cg", tol=1e-12, maxiter=2000,
                                           preconditioner="jacobi", return_info=True)
    xo, oinfo, oit = osolve.solve_cg(A, b, rtol=1e-5, atol=1e-12, maxiter=2000, precond="jacobi")
    xr, rinfo, rit = restated_device(A, b, solver="cg", rtol=1e-5, atol=1e-12, maxiter=2000,
                                     precond="jacobi")
    print(f"drop-in cg: gpu {info['iters']} iters, restated {rit}, scipy {oit}")
    assert info["info"] == rinfo == 0 and info["iters"] == rit
    assert np.array_equal(x, xr)
    assert rel_l2(x, xo) < 1e-6
