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
    """scipy BiCGSTAB (or CG, GMRES) on b and on nrep-1 copies of b perturbed
    by one ulp per entry: the reference's own rounding-noise floor.  Returns
    (iteration counts, infos, x of the unperturbed run, max relative L2
    spread of the perturbed solutions from it)."""
    from oracle import solve as osolve
    run = {"cg": osolve.solve_cg, "gmres": osolve.solve_gmres}.get(solver, osolve.solve_bicgstab)
    its, infos, xs = [], [], []
    for k in range(nrep):
        bb = b if k == 0 else b * (1 + np.random.default_rng(k).choice([-1, 1], len(b)) * 2.0 ** -52)
        x, info, it = run(A, bb, **kw)
        its.append(it)
        infos.append(info)
        xs.append(x)
    spread = max(rel_l2(x, xs[0]) for x in xs[1:]) if nrep > 1 else 0.0
    return its, infos, xs[0], spread


def _dot_fsum(a, b):
    import math
    return math.fsum((a * b).tolist())


def _dot_reversed(a, b):
    return float(np.dot(a[::-1], b[::-1]))


def _dot_chunked(a, b, chunk=256):
    p = a * b
    n = len(p)
    pad = (-n) % chunk
    if pad:
        p = np.concatenate([p, np.zeros(pad)])
    return float(np.sum(np.sum(p.reshape(-1, chunk), axis=1)))


DOT_ORDERS = {"fsum": _dot_fsum, "reversed": _dot_reversed, "chunked": _dot_chunked}


def _butterfly64(v):
    """wave64 xor-shuffle sum (v += shfl_xor(v, o), o = 32..1): lane 0's value."""
    idx = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        v = v + v[..., idx ^ o]
    return v[..., 0]


def _block_sum256(vals):
    """256-lane workgroup sum: each wave's butterfly, then waves 0..3 in order."""
    ws = _butterfly64(vals.reshape(vals.shape[0], 4, 64))
    s = ws[:, 0]
    for q in range(1, 4):
        s = s + ws[:, q]
    return s


def dot_device_order(cus):
    """The device GMRES's dot product order: G = min(ceil(n/256), 2048,
    8 cus) workgroups of 256 lanes, each lane summing its grid-stride
    products in order, workgroup sums as _block_sum256, then the G partials
    summed the same way by one workgroup (lane j taking partials j, j+256,
    ...)."""
    def dot(a, b):
        n = len(a)
        G = max(1, min((n + 255) // 256, 2048, 8 * cus))
        p = a * b
        T = G * 256
        acc = np.zeros(T)
        for start in range(0, n, T):
            c = p[start:start + T]
            acc[:len(c)] = acc[:len(c)] + c
        part = _block_sum256(acc.reshape(G, 256))
        t = np.zeros(256)
        for start in range(0, G, 256):
            c = part[start:start + 256]
            t[:len(c)] = t[:len(c)] + c
        return float(_block_sum256(t.reshape(1, 256))[0])
    return dot


def gmres_band(A, b, **kw):
    """scipy GMRES, and the restated GMRES (oracle.solve.gmres_restated)
    with the dot products summed in three other orders (exactly rounded,
    reversed, in chunks of 256) and the x update summed in k order: GMRES's
    own summation-order noise, which is the only arithmetic the device's
    solve does in another order.  Returns (inner iteration counts, cycle
    counts, infos, scipy's x, max relative L2 spread of the others from
    it)."""
    from oracle import solve as osolve
    x0, info0, it0, _, cyc0 = osolve.solve_gmres(A, b, counts=True, **kw)
    its, cycs, infos, spread = [it0], [cyc0], [info0], 0.0
    rk = {k: v for k, v in kw.items() if k in ("x0", "rtol", "atol", "maxiter", "precond", "restart")}
    if rk.get("restart") is None:
        rk["restart"] = 20
    for dot in DOT_ORDERS.values():
        x, info, it, cyc = osolve.gmres_restated(A, b, dot=dot, xupdate="sequential", **rk)
        its.append(it)
        cycs.append(cyc)
        infos.append(info)
        spread = max(spread, rel_l2(x, x0))
    return its, cycs, infos, x0, spread
