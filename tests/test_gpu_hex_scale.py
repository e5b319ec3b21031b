"""Structured hexes at full size, 128^3 = 2,097,152 cells.

* The momentum system (diffusion, upwind convection by a divergence-free
  flux, Dirichlet walls, source and the van Leer MUSCL correction from
  Green-Gauss gradients): device assembly equal to the oracle's (structure
  exact, values and right-hand side within 1e-14), then BiCGSTAB + ILU(0)
  with the FP64 true residual checked on the host.
* The SIMPLE pressure-correction system: FP64 and mixed precision (FP32
  SpMV / preconditioner / Krylov vectors inside FP64 reliable updates) both
  to relative residual 1e-10, each verified in FP64 on the host with the
  device-assembled matrix, and the mixed solution's error against the FP64
  one reported.  (The oracle's per-face pressure assembly is a Python loop,
  too slow at 6.3M faces; the pressure coefficients are checked against it
  at smaller sizes.)"""

import numpy as np
import pytest

from helpers import oracle_system, rel_l2
from oracle import solve as osolve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cases():
    from kcfd_amd import workloads
    return workloads.hex_momentum(128, seed=1002), workloads.hex_pressure(128, seed=1002)


def test_hex128_momentum(ctx, cases):
    from kcfd_amd import device, workloads
    case = cases[0]
    print("hex 128^3 momentum: device assembly vs the oracle", flush=True)
    dm = device.DeviceMesh(case.mesh, ctx)
    A = workloads.assemble_device(dm, case)
    Ao, bo, _ = oracle_system(case)
    Ag = A.to_scipy()
    assert np.array_equal(Ag.indptr, Ao.indptr) and np.array_equal(Ag.indices, Ao.indices)
    ev = np.max(np.abs(Ag.data - Ao.data) / np.maximum(np.abs(Ao.data), 1e-300))
    b = A.get_rhs()
    eb = np.max(np.abs(b - bo)) / np.max(np.abs(bo))
    A.setup_precond("ilu0", "fp64", order="color")
    x, res = A.solve(b, rtol=1e-8, precond="ilu0")
    tr = osolve.true_rel_residual(Ao, x, bo)
    print(f"hex 128^3 momentum: nnz {Ao.nnz}, values max rel {ev:.2e}, rhs {eb:.2e}; BiCGSTAB+ILU(0) "
          f"{res.iters} it, resid gpu {res.rel_resid:.2e} cpu {tr:.2e}")
    assert ev <= 1e-14 and eb <= 1e-14
    assert res.info == 0 and tr < 1e-8


def test_hex128_pressure_mixed_vs_fp64_to_1e10(ctx, cases):
    from kcfd_amd import device, workloads
    case = cases[1]
    print("hex 128^3 pressure: FP64 and mixed to 1e-10", flush=True)
    dm = device.DeviceMesh(case.mesh, ctx)
    A = workloads.assemble_device(dm, case)
    Ah = A.to_scipy()
    b = case.rhs
    out = {}
    for prec in ("fp64", "mixed"):
        A.setup_precond("ilu0", prec, order="color")
        x, res = A.solve(b, rtol=1e-10, precond="ilu0", precision=prec)
        tr = osolve.true_rel_residual(Ah, x, b)
        out[prec] = (x, res, tr)
        print(f"  {prec}: {res.iters} it, {res.refinements} replacements, info {res.info}, resid gpu "
              f"{res.rel_resid:.2e} cpu {tr:.2e}, {res.ms:.1f} ms", flush=True)
        assert res.info == 0 and tr < 1e-10
        assert abs(tr - res.rel_resid) <= 1e-3 * tr + 1e-16
    err = rel_l2(out["mixed"][0], out["fp64"][0])
    print(f"  mixed vs FP64 solution: relative L2 {err:.2e}")
    assert err < 1e-6
