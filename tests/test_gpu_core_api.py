"""The drop-in ``core`` API on the GPU, used the way the reference's
verification scripts use ``core`` (verification/case16_preconditioner.py,
case6_muscl.py): fields and a system are built with the operator functions,
then ``solve_linear_system`` / ``create_preconditioner`` are called with the
reference's arguments.  Everything is compared with the CPU oracle."""

import numpy as np
import pytest
from scipy import sparse

from helpers import rel_l2, restated_device
from oracle import fvm as ofvm, golden, solve as osolve

pytestmark = pytest.mark.gpu


def _channel_mesh(nx=80, ny=80):
    from kcfd_amd.mesh import build_fvmesh_from_arrays
    nodes, cells, bf = golden.channel_arrays(1.0, 1.0, nx, ny)
    return build_fvmesh_from_arrays(nodes, np.array(cells), {k: np.array(v) for k, v in bf.items()},
                                    ndim=2)


def _case16_through_core(mesh, gamma=0.01, u=1.0):
    from kcfd_amd import core
    system = core.FVMSystem(mesh.n_cells)
    g = core.ScalarField(mesh, "gamma", default=gamma)
    F = u * mesh.face_area * mesh.face_normal[:, 0]
    phi = core.ScalarField(mesh, "phi")
    phi.set_boundary("inlet", 1.0)
    phi.set_boundary("wall_bottom", 0.0)
    phi.set_boundary("wall_top", 0.0)
    bc = {"inlet": {"type": "dirichlet"}, "wall_bottom": {"type": "dirichlet"},
          "wall_top": {"type": "dirichlet"}, "outlet": {"type": "zero_gradient"}}
    core.diffusion_operator(mesh, g, system)
    core.convection_operator_upwind(mesh, F, system)
    core.apply_boundary_conditions(mesh, phi, g, F, system, bc)
    return system


def test_case16_assembly_and_solve(ctx):
    """Case 16 assembled by the operators equals the reference's inline
    assembly (golden.case16_system); the Jacobi solve at the figure's
    settings equals the restated scipy solver in the device's dot order
    (figure: 95)."""
    from kcfd_amd import core
    mesh = _channel_mesh()
    system = _case16_through_core(mesh)
    A = system.to_sparse()
    Ao, bo = golden.case16_system()
    A.sort_indices()
    Ao = sparse.csr_matrix(Ao)
    Ao.sort_indices()
    assert np.array_equal(A.indptr, Ao.indptr) and np.array_equal(A.indices, Ao.indices)
    assert np.max(np.abs(A.data - Ao.data)) <= 1e-14 * np.max(np.abs(Ao.data))
    assert np.max(np.abs(system.rhs - bo)) <= 1e-14 * np.max(np.abs(bo))
    x, info = core.solve_linear_system(system, method="bicgstab", tol=1e-10, maxiter=2000,
                                       preconditioner="jacobi", rtol=0.0, return_info=True)
    # the system the API solved: the operators' own matrix and rhs (equal to
    # the inline assembly to 1e-14, not necessarily bit for bit)
    xr, rinfo, rit, _ = restated_device(A, np.asarray(system.rhs, np.float64), rtol=0.0, atol=1e-10,
                                        maxiter=2000, precond="jacobi")
    print(f"case16 via core: iters={info['iters']} info={info['info']}; restated (device order) {rit} "
          f"info {rinfo} (figure {golden.CASE16_ITERS['jacobi']})")
    assert (info["info"], info["iters"]) == (rinfo, rit)
    assert np.array_equal(x, xr)


def test_conv_diff_muscl_gradient_and_mixed_solve(ctx):
    """Config-1-type system with a MUSCL deferred correction through the API:
    GG gradient and MUSCL correction equal the oracle's; the mixed-precision
    colour-ILU(0) solve is FP64-verified below 1e-8 and agrees with scipy."""
    from kcfd_amd import core, workloads
    case = workloads.quad_convdiff(48, seed=1001)
    mesh = case.mesh
    rec = dict(case.recipe)
    gamma, F = rec["diffusion"]["gamma"], rec["upwind"]["F"]
    bc = rec["boundary"]["bc"]
    xc = mesh.cell_center
    phi = core.ScalarField(mesh, "phi")
    phi.values[:] = np.sin(3 * xc[:, 0]) * np.cos(2 * xc[:, 1])
    for nm, f in mesh.boundary_patches.items():
        phi.set_boundary(nm, np.zeros(len(f)))
    grad = core.green_gauss_gradient(phi)
    grad_o = ofvm.green_gauss_gradient(mesh, phi.values, phi.boundary_values)
    assert np.max(np.abs(grad[:, :2] - grad_o[:, :2])) <= 1e-13 * max(1.0, np.max(np.abs(grad_o)))
    lsq = core.least_squares_gradient(phi)
    lsq_o = ofvm.least_squares_gradient(mesh, phi.values, phi.boundary_values)
    assert np.max(np.abs(lsq - lsq_o)) <= 1e-12 * max(1.0, np.max(np.abs(lsq_o)))
    corr = core.muscl_deferred_correction(mesh, phi, F, grad, "van_leer")
    corr_o = ofvm.muscl_deferred_correction(mesh, phi.values, F, grad_o, "van_leer")
    assert np.max(np.abs(corr - corr_o)) <= 1e-13 * max(1.0, np.max(np.abs(corr_o)))

    system = core.FVMSystem(mesh.n_cells)
    core.diffusion_operator(mesh, gamma, system)
    core.convection_operator_upwind(mesh, F, system)
    core.apply_boundary_conditions(mesh, phi, gamma, F, system, bc)
    system.rhs = system.rhs + corr                # case6_muscl.py:117
    A = system.to_sparse()
    b = system.rhs.copy()
    x, info = core.solve_linear_system(system, method="bicgstab", tol=0.0, maxiter=5000,
                                       preconditioner="ilu0", rtol=1e-8, precision="mixed",
                                       order="color", return_info=True)
    res = osolve.true_rel_residual(A, x, b)
    xo, info_o, it_o = osolve.solve_bicgstab(A, b, rtol=1e-12, precond="ilu0_color")
    print(f"mixed colour ILU via core: iters={info['iters']} refinements={info['refinements']} "
          f"fallback={info['fallback']} resid={res:.2e} err vs fp64={rel_l2(x, xo):.2e}")
    assert info["info"] == 0 and res < 1e-8 and info["fallback"] == 0
    assert rel_l2(x, xo) < 1e-6


def test_create_preconditioner_colour_ilu(ctx):
    """create_preconditioner(A, 'ilu0', order='color') returns (M, info) with
    M.matvec = the oracle's ILU(0) of P A P^T."""
    from kcfd_amd import core, workloads
    from helpers import oracle_system
    A, b, _ = oracle_system(workloads.hex_pressure(10, seed=1002))
    M, info = core.create_preconditioner(A, "ilu0", order="color")
    assert info["method"] == "ilu0"
    z = np.random.default_rng(3).uniform(-1, 1, A.shape[0])
    ref = osolve.ILU0Ordered(A, osolve.color_order(A)[0])
    assert rel_l2(M.matvec(z), ref.solve(z)) <= 1e-13
    with pytest.raises(ValueError):
        core.create_preconditioner(A, "nonsense")


def test_benchmark_solvers(ctx):
    """core/gpu_solver.py:202-250: the same keys, the HIP backend, finite
    median times, x_direct = scipy's spsolve."""
    import math
    from scipy.sparse.linalg import spsolve
    from helpers import oracle_system
    from kcfd_amd import core, workloads
    A, b, _ = oracle_system(workloads.quad_convdiff(24, seed=1001))
    r = core.benchmark_solvers(A, b, n_runs=2)
    assert set(r) == {"backend", "cpu_direct", "x_direct", "cpu_bicgstab", "gpu_bicgstab"}
    assert r["backend"] == "hip"
    assert all(math.isfinite(r[k]) and r[k] > 0 for k in ("cpu_direct", "cpu_bicgstab", "gpu_bicgstab"))
    assert np.array_equal(r["x_direct"], spsolve(A, b))
