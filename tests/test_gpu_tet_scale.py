"""Unstructured tetrahedra at full size: Delaunay of ~200k jittered points
(about 1.27M tets, cells randomly permuted then RCM-reordered), two-fluid
momentum with Schiller-Naumann drag assembled on the device, Jacobi
BiCGSTAB in FP64 and mixed precision."""

import numpy as np
import pytest

from helpers import oracle_system, rel_l2
from oracle import solve as osolve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tet_case():
    from kcfd_amd import workloads
    return workloads.tet_twofluid(200_000, seed=1003)


def test_tet_twofluid_full_size(ctx, tet_case):
    from kcfd_amd import device, workloads
    case = tet_case
    print(f"tets {case.mesh.n_cells}, faces {case.mesh.n_faces}")
    assert case.mesh.n_cells > 1_000_000
    A, b, diag = oracle_system(case)
    dm = device.DeviceMesh(case.mesh, ctx)
    M = workloads.assemble_device(dm, case)
    Ad = M.to_scipy()
    Ad.sort_indices()
    assert np.array_equal(Ad.indptr, A.indptr) and np.array_equal(Ad.indices, A.indices)
    err = np.max(np.abs(Ad.data - A.data) / np.maximum(np.abs(A.data), 1e-300))
    print(f"assembly: nnz {A.nnz}, max rel diff {err:.2e}")
    assert err <= 1e-14
    xo, info_o, it_o = osolve.solve_bicgstab(A, b, rtol=1e-8, precond="jacobi")
    x, res = M.solve(b, rtol=1e-8, precond="jacobi")
    print(f"FP64 jacobi: gpu iters={res.iters} scipy {it_o}; relL2 {rel_l2(x, xo):.2e}")
    assert res.info == 0 and abs(res.iters - it_o) <= 2 and rel_l2(x, xo) <= 1e-10
    xm, rm = M.solve(b, rtol=1e-8, precond="jacobi", precision="mixed")
    tr = osolve.true_rel_residual(A, xm, b)
    print(f"mixed jacobi: iters={rm.iters} refinements={rm.refinements} resid={tr:.2e} "
          f"err vs fp64 {rel_l2(xm, xo):.2e}")
    assert rm.info == 0 and tr < 1e-8
    M.close()
    dm.close()


def test_tet_two_fluid_pressure_step_full_size(ctx, tet_case):
    """The two-fluid pressure-correction step (models/two_fluid.py:641-725)
    on the full tet mesh: matrix and right-hand side against the oracle's
    (vectorised, same COO order) to 1e-14, p' FP64-verified to 1e-10 with
    the oracle's matrix (spsolve does not fit this size), and the velocity /
    pressure corrections against the oracle's correction of the device p'."""
    from kcfd_amd import device, simple
    from kcfd_amd._lib import DevBuf
    from oracle import fvm as ofvm
    mesh = tet_case.mesh
    rng = np.random.default_rng(31)
    n, nf = mesh.n_cells, mesh.n_faces
    al = rng.uniform(0.05, 0.95, n)
    f = dict(alpha_l=al, alpha_g=1.0 - al, aP_l=np.exp(rng.normal(0, 0.5, n)),
             aP_g=0.1 * np.exp(rng.normal(0, 0.5, n)), mf_l=rng.normal(0, 0.1, nf),
             mf_g=rng.normal(0, 0.01, nf))
    Ul, Ug, p = rng.normal(0, 1, (n, 3)), rng.normal(0, 1, (n, 3)), rng.normal(0, 1, n)
    A_o, b_o = ofvm.two_fluid_pressure_system(mesh, f["alpha_l"], f["alpha_g"], f["aP_l"], f["aP_g"],
                                              998.2, 1.225, f["mf_l"], f["mf_g"])
    dm = device.DeviceMesh(mesh, ctx)
    d = {k: DevBuf.from_array(ctx, v) for k, v in f.items()}
    Uld, Ugd, pdv = (DevBuf.from_array(ctx, Ul.reshape(-1)), DevBuf.from_array(ctx, Ug.reshape(-1)),
                     DevBuf.from_array(ctx, p))
    res = simple.two_fluid_pressure_correction_step(dm, d["alpha_l"], d["alpha_g"], d["aP_l"], d["aP_g"],
                                                    998.2, 1.225, d["mf_l"], d["mf_g"], Uld, Ugd, pdv, 0.3,
                                                    precond="jacobi", rtol=1e-10)
    Ag = res.matrix.to_scipy()
    assert np.array_equal(Ag.indptr, A_o.indptr) and np.array_equal(Ag.indices, A_o.indices)
    ea = np.max(np.abs(Ag.data - A_o.data) / np.abs(A_o.data))
    eb = np.max(np.abs(res.matrix.get_rhs() - b_o)) / np.max(np.abs(b_o))
    pp = res.pp.download()
    tr = osolve.true_rel_residual(A_o, pp, b_o)
    Ul_o, Ug_o, p_o, _ = ofvm.two_fluid_correction(mesh, pp, f["alpha_l"], f["alpha_g"], f["aP_l"],
                                                   f["aP_g"], Ul, Ug, p, 0.3)
    eu = max(np.max(np.abs(Uld.download().reshape(n, 3) - Ul_o)) / np.max(np.abs(Ul_o)),
             np.max(np.abs(Ugd.download().reshape(n, 3) - Ug_o)) / np.max(np.abs(Ug_o)))
    ep = np.max(np.abs(pdv.download() - p_o)) / np.max(np.abs(p_o))
    print(f"two-fluid step on {n} tets: A {ea:.1e} b {eb:.1e}; Jacobi BiCGSTAB {res.solve['iters']} it, "
          f"p' resid (FP64, host) {tr:.1e}; corrections U {eu:.1e} p {ep:.1e}")
    assert ea <= 1e-14 and eb <= 1e-14
    assert res.solve["info"] == 0 and tr < 1e-10
    assert eu <= 1e-13 and ep <= 1e-15
    dm.close()
