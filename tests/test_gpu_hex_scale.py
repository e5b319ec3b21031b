"""Structured hexes at full size, 128^3 = 2,097,152 cells.

* The momentum system (diffusion, upwind convection by a divergence-free
  flux, Dirichlet walls, source and the van Leer MUSCL correction from
  Green-Gauss gradients): device assembly equal to the oracle's (structure
  exact, values and right-hand side within 1e-14), then BiCGSTAB + ILU(0)
  with the FP64 true residual checked on the host.
* The SIMPLE pressure-correction system of the benchmark (config 2): the
  device assembly equal to the oracle's (structure exact, values within
  1e-14); the FP64 BiCGSTAB + colour-ordered ILU(0) solve to 1e-8 equal bit
  for bit to the restated scipy solver in the device's dot order (its
  iterations, info and the SHA-256 of x are a fixture,
  tests/golden/restated_hex128.json, written by
  tests/golden/make_restated_fixtures.py: the restatement takes two minutes
  on a CPU); FP64 and mixed precision (FP32 SpMV / preconditioner / Krylov
  vectors inside FP64 reliable updates) both to relative residual 1e-10,
  each verified in FP64 on the host, and the mixed solution's error against
  the FP64 one reported."""

import hashlib
import json
import os

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


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<f8").tobytes()).hexdigest()


def test_hex128_pressure_assembly_and_fp64_solve(ctx, cases):
    from kcfd_amd import device, workloads
    case = cases[1]
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "restated_hex128.json")))
    fx = fx["hex128_pressure_fp64_ilu0_color_rtol1e-8_rows2"]
    print("hex 128^3 pressure: device assembly vs the oracle", flush=True)
    dm = device.DeviceMesh(case.mesh, ctx)
    Ad = workloads.assemble_device(dm, case)
    Ao, bo, diag_o = oracle_system(case)
    assert _sha(Ao.data) == fx["A_sha256"] and _sha(bo) == fx["b_sha256"]
    Ag = Ad.to_scipy()
    assert np.array_equal(Ag.indptr, Ao.indptr) and np.array_equal(Ag.indices, Ao.indices)
    ev = np.max(np.abs(Ag.data - Ao.data) / np.maximum(np.abs(Ao.data), 1e-300))
    et = np.max(np.abs(Ad.get_diag_tracker() - diag_o) / np.maximum(np.abs(diag_o), 1e-300))
    print(f"  nnz {Ao.nnz}: values max rel {ev:.2e} (bit-equal {np.array_equal(Ag.data, Ao.data)}), "
          f"tracker {et:.2e}", flush=True)
    assert ev <= 1e-14 and et <= 1e-14
    M = device.DeviceMatrix.from_csr(ctx, Ao)
    M.setup_precond("ilu0", "fp64", order="color")
    x, res = M.solve(bo, rtol=1e-8, precond="ilu0", rows_per_thread=2)
    print(f"  FP64 BiCGSTAB + colour ILU(0): gpu {res.iters} it info {res.info} half {res.half_step}, "
          f"resid {res.rel_resid:.2e}, {res.ms:.1f} ms | restated (device order) {fx['iters']} it "
          f"info {fx['info']}", flush=True)
    assert (res.info, res.iters, res.half_step) == (fx["info"], fx["iters"], fx["half_step"])
    assert _sha(x) == fx["x_sha256"]


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
