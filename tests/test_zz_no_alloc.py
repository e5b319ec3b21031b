"""kcfd_solve_noalloc against kcfd_solve, through the C-ABI.

The allocation-free entry point runs the same kernels on a workspace in a
static device array.  Two calls in a row, the second on a smaller system,
each compared with kcfd_solve on the same system: x bit for bit, and the
same info, iteration count, half-step flag and reliable-update count --
for FP64 BiCGSTAB with Jacobi and with the colour- and segment-ordered
ILU(0), mixed precision and CG.  A system whose workspace does not fit the array is
refused with a message.
"""

import ctypes

import numpy as np
import pytest
from scipy import sparse

from helpers import oracle_system

pytestmark = pytest.mark.gpu


def _both(ctx, A, b, pc, precision="fp64", method="bicgstab", order=None):
    from kcfd_amd import _lib as L, device
    from kcfd_amd._lib import DevBuf
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond(pc, precision, order=order)
    opts = device.default_opts(rtol=1e-8, precond=pc, precision=precision, method=method)
    out = []
    for fn in (L.lib().kcfd_solve, L.lib().kcfd_solve_noalloc):
        bd = DevBuf.from_array(ctx, b)
        xd = DevBuf.from_array(ctx, np.zeros(A.shape[0]))
        res = L.SolveResult()
        L.check(fn(M.h, bd.p, xd.p, ctypes.byref(opts), ctypes.byref(res)))
        out.append((xd.download(), res))
    M.close()
    return out


@pytest.mark.parametrize("kind", ["jacobi", "ilu0_color", "ilu0_segment", "mixed", "cg"])
def test_noalloc_equals_solve(ctx, kind):
    from kcfd_amd import workloads
    if kind == "cg":
        cases = [workloads.hex_pressure(24, seed=1002), workloads.hex_pressure(12, seed=1003)]
    elif kind in ("ilu0_color", "ilu0_segment"):
        cases = [workloads.hex_pressure(32, seed=1002), workloads.quad_convdiff(40, seed=1004)]
    else:
        cases = [workloads.quad_convdiff(96, seed=1001), workloads.quad_convdiff(48, seed=1005)]
    for call, case in enumerate(cases):
        A, b, _ = oracle_system(case)
        if kind == "jacobi":
            (xa, ra), (xn, rn) = _both(ctx, A, b, "jacobi")
        elif kind in ("ilu0_color", "ilu0_segment"):
            (xa, ra), (xn, rn) = _both(ctx, A, b, "ilu0", order=kind[5:])
        elif kind == "mixed":
            (xa, ra), (xn, rn) = _both(ctx, A, b, "jacobi", precision="mixed")
        else:
            (xa, ra), (xn, rn) = _both(ctx, A, b, "jacobi", method="cg")
        print(f"{kind} call {call}: n={A.shape[0]} kcfd_solve {ra.iters} it info {ra.info}; "
              f"noalloc {rn.iters} it info {rn.info}", flush=True)
        assert (rn.info, rn.iters, rn.half_step, rn.refinements) == \
            (ra.info, ra.iters, ra.half_step, ra.refinements)
        assert np.array_equal(xn, xa)
        assert ra.info == 0


def test_noalloc_refuses_oversize(ctx):
    from kcfd_amd import _lib as L, device
    from kcfd_amd._lib import DevBuf
    n = 1_600_000   # Jacobi FP64 needs ~77 MB of workspace, the array holds 64 MiB
    A = sparse.diags(np.full(n, 2.0)).tocsr()
    M = device.DeviceMatrix.from_csr(ctx, A)
    M.setup_precond("jacobi", "fp64")
    bd, xd = DevBuf.from_array(ctx, np.ones(n)), DevBuf.from_array(ctx, np.zeros(n))
    opts = device.default_opts(rtol=1e-8, precond="jacobi")
    res = L.SolveResult()
    rc = L.lib().kcfd_solve_noalloc(M.h, bd.p, xd.p, ctypes.byref(opts), ctypes.byref(res))
    assert rc == L.KCFD_EARG
    assert "scratch array" in L.lib().kcfd_last_error().decode()
    M.close()
