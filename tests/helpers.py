"""Test helpers: replay a workload recipe through the CPU oracle."""

from __future__ import annotations

import numpy as np

from oracle import fvm as ofvm


def oracle_system(case):
    """Assemble ``case`` with the oracle; returns (A csr, rhs, diag tracker)."""
    mesh = case.mesh
    s = ofvm.OSystem(mesh.n_cells)
    F_cur = None
    for op, kw in case.recipe:
        if op == "mass_flux":
            F_cur = ofvm.compute_mass_flux(mesh, kw["U"], kw["Ub"], kw["rho"])
        elif op == "diffusion":
            ofvm.diffusion_operator(mesh, kw["gamma"], s)
        elif op == "upwind":
            ofvm.convection_operator_upwind(mesh, kw["F"], s)
        elif op == "upwind_scaled":
            F_cur = F_cur * np.asarray(kw["scale_owner"])[mesh.owner]
            ofvm.convection_operator_upwind(mesh, F_cur, s)
        elif op == "boundary":
            ofvm.apply_boundary_conditions(mesh, kw["phib"], kw["gamma"], kw["F"], s, kw["bc"])
        elif op == "boundary_scaled":
            ofvm.apply_boundary_conditions(mesh, kw["phib"], kw["gamma"], F_cur, s, kw["bc"])
        elif op == "pressure":
            pressure_correction(mesh, kw["dP"], kw["rho"], kw["dirichlet"], s)
        elif op == "source":
            ofvm.source_term(mesh, kw["S"], s)
        elif op == "cell_terms":
            idx = np.arange(mesh.n_cells)
            k = np.asarray(kw["kd"]) * mesh.cell_volume
            s.add_entries(idx, idx, k)
            s.add_source(idx, k * np.asarray(kw["ks"]))
        elif op == "temporal":
            ofvm.temporal_operator(mesh, kw["rho"], kw["dt"], kw["phi_old"], s)
        elif op == "muscl":
            g = ofvm.green_gauss_gradient(mesh, kw["phi"], kw["phib"])
            c = ofvm.muscl_deferred_correction(mesh, kw["phi"], kw["F"], g, kw["limiter"])
            s.rhs = s.rhs + c
        else:
            raise ValueError(op)
    A = s.to_sparse()
    rhs = s.rhs.copy() if case.rhs is None else np.asarray(case.rhs, float).copy()
    return A, rhs, s.diag.copy()


def pressure_correction(mesh, dP, rho, dirichlet, s):
    """The oracle's pressure-correction assembly (oracle/fvm.py)."""
    ofvm.pressure_correction_operator(mesh, dP, rho, dirichlet, s)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def oracle_band(A, b, nrep=6, solver="bicgstab", **kw):
    """scipy BiCGSTAB (or CG) on b and on nrep-1 copies of b perturbed by one
    ulp per entry: the reference's own rounding-noise floor.  Returns
    (iteration counts, infos, x of the unperturbed run, max relative L2
    spread of the perturbed solutions from it)."""
    from oracle import solve as osolve
    run = osolve.solve_cg if solver == "cg" else osolve.solve_bicgstab
    its, infos, xs = [], [], []
    for k in range(nrep):
        bb = b if k == 0 else b * (1 + np.random.default_rng(k).choice([-1, 1], len(b)) * 2.0 ** -52)
        x, info, it = run(A, bb, **kw)
        its.append(it)
        infos.append(info)
        xs.append(x)
    spread = max(rel_l2(x, xs[0]) for x in xs[1:]) if nrep > 1 else 0.0
    return its, infos, xs[0], spread
