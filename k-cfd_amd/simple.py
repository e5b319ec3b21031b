"""The SIMPLE pressure-correction steps on the device, around the hot-path
solve: single-phase (models/single_phase.py:274-365) and two-fluid
(models/two_fluid.py:641-725).

``pressure_correction_step`` keeps every array on the device: d_P from the
momentum diagonals (kcfd_simple_dp), the pressure-correction matrix
(kcfd_asm_pressure) with the mass-imbalance right-hand side
(kcfd_simple_rhs) and, without a Dirichlet p' patch, the reference-cell row
(kcfd_simple_fix_reference), the solve of A p' = b (the reference calls
spsolve; here BiCGSTAB or CG to a tight tolerance), and the fused velocity
and pressure correction (kcfd_simple_correct).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import device
from ._lib import DevBuf, check, lib


@dataclass
class StepResult:
    pp: DevBuf          # p' (device, n)
    dP: DevBuf          # d_P (device, n)
    solve: dict         # the solve's result
    matrix: device.DeviceMatrix
    ratio: float = 0.0  # ||p'|| / max(||p|| + 1e-10, 1e-10) after the update (the reference's return value)


def pressure_correction_step(dm: device.DeviceMesh, aP, U: DevBuf, p: DevBuf, mf: DevBuf, rho: float,
                             alpha_p: float, dirichlet, precond="ilu0", order="color", rtol=1e-12,
                             method="bicgstab", matrix: device.DeviceMatrix | None = None) -> StepResult:
    """One pressure-correction step.  aP: ndim device buffers (the momentum
    diagonals); U (n*ndim, row-major), p (n) are updated in place; mf: face
    mass fluxes (n_faces); dirichlet: names of the Dirichlet p' patches."""
    n, nd = dm.n, dm.ndim
    pd = np.array([1 if nm in set(dirichlet) else 0 for nm in dm.patch_names], np.int32)
    dP = DevBuf(dm.ctx, n)
    check(lib().kcfd_simple_dp(dm.h, aP[0].p, aP[1].p, aP[2].p if nd == 3 else None, dP.p))
    A = matrix if matrix is not None else dm.new_system()
    check(lib().kcfd_asm_reset(A.h))
    check(lib().kcfd_asm_pressure(dm.h, A.h, dP.p, float(rho), pd.ctypes.data))
    check(lib().kcfd_simple_rhs(dm.h, A.h, mf.p))
    if not pd.any():
        check(lib().kcfd_simple_fix_reference(A.h))
    A.setup_precond(precond, "fp64", order=order if precond == "ilu0" else None)
    pp = DevBuf(dm.ctx, n)
    pp.fill(0.0)
    opts = device.default_opts(rtol=rtol, precond=precond, precision="fp64", method=method)
    res = A.solve_dev(A.rhs_ptr, pp.p, opts, setup=False)
    check(lib().kcfd_simple_correct(dm.h, pp.p, dP.p, pd.ctypes.data, float(alpha_p), U.p, p.p))
    # models/single_phase.py:376
    ratio = float(np.linalg.norm(pp.download()) / max(np.linalg.norm(p.download()) + 1e-10, 1e-10))
    return StepResult(pp, dP, res.as_dict(), A, ratio)


@dataclass
class TwoFluidStepResult:
    pp: DevBuf          # p' (device, n)
    solve: dict         # the solve's result
    matrix: device.DeviceMatrix
    ratio: float        # ||p'|| / max(||p||, 1e-15) after the update (the reference's return value)


def two_fluid_pressure_correction_step(dm: device.DeviceMesh, alpha_l: DevBuf, alpha_g: DevBuf,
                                       aP_l: DevBuf, aP_g: DevBuf, rho_l: float, rho_g: float,
                                       mf_l: DevBuf, mf_g: DevBuf, U_l: DevBuf, U_g: DevBuf, p: DevBuf,
                                       alpha_p: float, precond="ilu0", order="color", rtol=1e-12,
                                       method="bicgstab",
                                       matrix: device.DeviceMatrix | None = None) -> TwoFluidStepResult:
    """``TwoFluidSolver._solve_pressure_correction_two_phase``
    (models/two_fluid.py:641-725) on the device: the two-phase
    pressure-correction system (kcfd_twofluid_pressure: coefficients, mass
    imbalance, 1e10 on the first diagonal), the solve of A p' = b (the
    reference calls solve_pressure_correction, i.e. spsolve; here the device
    Krylov solver to ``rtol``), a p' with NaNs replaced by zeros as in the
    reference, and the velocity corrections of both phases with
    p += alpha_p p' (kcfd_twofluid_correct).  U_l, U_g (n*ndim, row-major)
    and p are updated in place."""
    n = dm.n
    A = matrix if matrix is not None else dm.new_system()
    check(lib().kcfd_asm_reset(A.h))
    check(lib().kcfd_twofluid_pressure(dm.h, A.h, alpha_l.p, alpha_g.p, aP_l.p, aP_g.p, float(rho_l),
                                       float(rho_g), mf_l.p, mf_g.p))
    A.setup_precond(precond, "fp64", order=order if precond == "ilu0" else None)
    pp = DevBuf(dm.ctx, n)
    pp.fill(0.0)
    opts = device.default_opts(rtol=rtol, precond=precond, precision="fp64", method=method)
    res = A.solve_dev(A.rhs_ptr, pp.p, opts, setup=False)
    ph = pp.download()
    if np.any(np.isnan(ph)):           # the reference's NaN guard (:694-695)
        pp.fill(0.0)
        ph[:] = 0.0
    check(lib().kcfd_twofluid_correct(dm.h, pp.p, alpha_l.p, alpha_g.p, aP_l.p, aP_g.p, float(alpha_p),
                                      U_l.p, U_g.p, p.p))
    ratio = float(np.linalg.norm(ph) / max(np.linalg.norm(p.download()), 1e-15))
    return TwoFluidStepResult(pp, res.as_dict(), A, ratio)


__all__ = ["pressure_correction_step", "StepResult", "two_fluid_pressure_correction_step",
           "TwoFluidStepResult"]
