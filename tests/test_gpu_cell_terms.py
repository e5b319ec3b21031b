"""GPU parity of the operators a reference caller strings together around
the convection-diffusion core, through the drop-in ``core`` API:
temporal_operator (core/fvm_operators.py:132-148), source_term (:151-161),
linearized_source (:164-175), apply_boundary_conditions with Dirichlet,
Neumann and zero-gradient patches (:178-225), a caller's in-place edit of
the diagonal tracker (the floor of models/two_fluid.py:606-609) and then
under_relax, which reads that tracker (:228-240).  The device system
(matrix, right-hand side and tracker) equals the oracle's: structure exact,
values within 1e-14 of the sum of the absolute contributions to each
entry (the device sums an entry's contributions in another order than the
reference's COO duplicates; where they cancel -- 400-fold on the tet
mesh -- a bound relative to the value itself would be ill-conditioned),
right-hand side and tracker within 1e-14 relative.  Also the
Schiller-Naumann drag closure
(models/closure.py:13-34, 72-92) on the device against the oracle on both
Re branches and at both floors.
"""

import numpy as np
import pytest
from scipy import sparse

from oracle import fvm as ofvm

pytestmark = pytest.mark.gpu


def _mesh(which):
    from kcfd_amd import workloads
    return {"quad": lambda: workloads.quad_mesh(12),
            "hex": lambda: workloads.hex_mesh(6),
            "tet": lambda: workloads.tet_mesh(400, seed=1003)}[which]()


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if len(b) else 0.0


@pytest.mark.parametrize("which", ["quad", "hex", "tet"])
@pytest.mark.parametrize("alpha", [0.7, 1.0])
def test_operator_chain_with_tracker_edit(ctx, which, alpha):
    from kcfd_amd import core
    mesh = _mesh(which)
    n, nf = mesh.n_cells, mesh.n_faces
    rng = np.random.default_rng(17)
    gamma = np.exp(rng.normal(-3, 0.5, n))
    F = rng.normal(0, 0.05, nf)
    phi_old = rng.normal(0, 1, n)
    S = rng.normal(0, 1, n)
    # some cells get Sp large enough to drive the tracker below 1e-20, so the
    # caller's floor changes what under_relax reads
    Sp = np.where(rng.uniform(size=n) < 0.2, 1e3, -np.abs(rng.normal(0, 1, n)))
    Su = rng.normal(0, 1, n)
    phi_vals = rng.normal(0, 1, n)
    names = list(mesh.boundary_patches)
    kinds = ["neumann", "dirichlet", "zero_gradient"]   # the tet mesh has one patch
    bc = {nm: {"type": kinds[i % 3], "value": 0.25 * (i + 1)} for i, nm in enumerate(names)}
    phib = {nm: rng.normal(0, 1, len(f)) for nm, f in mesh.boundary_patches.items()}
    assert any(b["type"] == "neumann" for b in bc.values())

    # device, through the drop-in API
    system = core.FVMSystem(n)
    phi = core.ScalarField(mesh, "phi")
    phi.values[:] = phi_vals
    for nm, v in phib.items():
        phi.set_boundary(nm, v)
    g = core.ScalarField(mesh, "gamma")
    g.values[:] = gamma
    core.diffusion_operator(mesh, g, system)
    core.convection_operator_upwind(mesh, F, system)
    core.temporal_operator(mesh, 2.5, 0.01, phi_old, system)
    core.source_term(mesh, S, system)
    core.linearized_source(mesh, Sp, Su, system)
    core.apply_boundary_conditions(mesh, phi, g, F, system, bc)
    d = system.diag                     # the caller edits the tracker in place
    low = d < 1e-20
    d[low] = 1e-10
    core.under_relax(system, phi, alpha)
    A = system.to_sparse()
    rhs, diag = system.rhs.copy(), system.diag.copy()

    # oracle
    o = ofvm.OSystem(n)
    ofvm.diffusion_operator(mesh, gamma, o)
    ofvm.convection_operator_upwind(mesh, F, o)
    ofvm.temporal_operator(mesh, 2.5, 0.01, phi_old, o)
    ofvm.source_term(mesh, S, o)
    ofvm.linearized_source(mesh, Sp, Su, o)
    ofvm.apply_boundary_conditions(mesh, phib, gamma, F, o, bc)
    o.diag[o.diag < 1e-20] = 1e-10
    ofvm.under_relax(o, phi_vals, alpha)
    Ao = o.to_sparse()
    Ao.sum_duplicates()
    A.sum_duplicates()
    r, c, v = o.coo()
    Aabs = sparse.coo_matrix((np.abs(v), (r, c)), shape=Ao.shape).tocsr()
    Aabs.sum_duplicates()
    ev = float(np.max(np.abs(A.data - Ao.data) / np.maximum(Aabs.data, 1e-300)))
    er, et = _rel(rhs, o.rhs), _rel(diag, o.diag)
    print(f"{which} alpha {alpha}: {low.sum()} tracker entries floored; values {ev:.1e} of the summed "
          f"|contributions| ({_rel(A.data, Ao.data):.1e} of the value), rhs {er:.1e} tracker {et:.1e}")
    assert low.sum() > 0
    assert np.array_equal(A.indptr, Ao.indptr) and np.array_equal(A.indices, Ao.indices)
    assert ev <= 1e-14 and er <= 1e-14 and et <= 1e-14


def test_neumann_closed_form(ctx):
    """A Neumann patch adds g A_f to the owner's rhs and nothing else
    (core/fvm_operators.py:214-217): on a uniform quad grid with only
    Neumann patches the rhs is g h on each boundary face, twice at corners,
    and the matrix is untouched."""
    from kcfd_amd import core
    mesh = _mesh("quad")
    n = mesh.n_cells
    system = core.FVMSystem(n)
    g = core.ScalarField(mesh, "gamma", default=0.3)
    phi = core.ScalarField(mesh, "phi")
    F = np.zeros(mesh.n_faces)
    bc = {nm: {"type": "neumann", "value": 2.0} for nm in mesh.boundary_patches}
    core.diffusion_operator(mesh, g, system)
    A0 = system.to_sparse().copy()
    core.apply_boundary_conditions(mesh, phi, g, F, system, bc)
    A1 = system.to_sparse()
    assert np.array_equal(A0.data, A1.data)
    h = 1.0 / 12
    expect = np.zeros(n)
    for f in mesh.boundary_patches.values():
        np.add.at(expect, mesh.owner[np.asarray(f)], 2.0 * h)
    assert np.allclose(system.rhs, expect, rtol=1e-14, atol=0)


def test_drag_closure(ctx):
    from kcfd_amd import closure
    rng = np.random.default_rng(5)
    n = 4000
    ul = rng.normal(0, 0.2, (n, 3))
    ug = ul + rng.normal(0, 1, (n, 3)) * np.exp(rng.uniform(-12, 1, (n, 1)))
    ug[:50] = ul[:50]                       # |u_rel| = 0: floored at 1e-15, Re floored at 1e-10
    alpha_g = rng.uniform(0.05, 0.6, n)
    rho_l, d_b, mu_l = 998.2, 0.005, 1.003e-3
    K = closure.drag_coefficient_implicit(alpha_g, rho_l, ug, ul, d_b, mu_l)
    Ko = ofvm.drag_coefficient_implicit(alpha_g, rho_l, ug, ul, d_b, mu_l)
    mag = np.maximum(np.sqrt(np.sum((ug - ul) ** 2, axis=1)), 1e-15)
    Re = rho_l * mag * d_b / mu_l
    print(f"drag: Re from {Re.min():.1e} to {Re.max():.1e}, {np.sum(Re >= 1000)} cells at C_D = 0.44; "
          f"max rel diff {_rel(K, Ko):.1e}")
    assert np.sum(Re < 1e-10) >= 50 and np.sum(Re >= 1000) > 100 and np.sum((Re > 1) & (Re < 1000)) > 100
    assert _rel(K, Ko) <= 1e-14
    # the C_D function alone: both branches, the switch and the floor
    Re2 = np.array([0.0, 1e-12, 1e-10, 0.5, 1.0, 100.0, 999.999, 1000.0, 1000.001, 5e4])
    cd = closure.schiller_naumann_drag(Re2)
    cdo = ofvm.schiller_naumann_drag(Re2)
    assert _rel(cd, cdo) <= 1e-14
    assert cd[-1] == cd[-2] == cd[-3] == 0.44
    assert abs(cd[6] - 24.0 / 999.999 * (1.0 + 0.15 * 999.999 ** 0.687)) <= 1e-14 * cd[6]   # 0.4383
    assert cd[0] == cd[1] == cd[2]
