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
