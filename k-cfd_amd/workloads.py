"""Seeded synthetic workloads (mesh + fields + an operator recipe).

Each generator returns a ``Case``: the mesh and an ordered recipe of the
reference operator calls with their inputs.  The recipe is played on the
device by ``assemble_device`` (the product path); tests replay the same
recipe through the CPU oracle, so both sides assemble the identical system.

Configurations (sizes are parameters; defaults are the full sizes):
  quad_convdiff  2D unit square quads, lognormal gamma (mu=ln 1e-2,
                 sigma=0.5), upwind convection by the curl of an 8-mode
                 stream function (exactly divergence-free face fluxes,
                 peak |u| at cell centres = 1), Dirichlet 0, b ~ U[-1,1].
  hex_pressure   3D unit cube hexes, SIMPLE pressure correction
                 (models/single_phase.py:288-325 form) with a_P ~
                 lognormal(0,1), dP = V/a_P, Dirichlet p'=0 on "outlet",
                 b ~ U[-1,1].
  hex_momentum   3D unit cube hexes, gamma = 1e-2, upwind convection by a
                 random divergence-free flux (circulation of an 8-mode
                 vector potential around each face), Dirichlet 0, rhs =
                 U[-1,1]*V plus the van Leer MUSCL deferred correction from
                 Green-Gauss gradients of a smooth field.
  hybrid_*       The hex_pressure / hex_momentum recipes on the mixed
                 hex/tet duct of mesh/hybrid_mesh_generator.py (2 x 0.1 x
                 0.1, hexes on the inlet half, 12 tets per hex beyond).
  tet_twofluid   Delaunay tetrahedra of jittered lattice points, cells
                 permuted then RCM-reordered; liquid-phase momentum of
                 models/two_fluid.py:520-612 (alpha_g ~ U[0.05,0.6],
                 diffusion alpha_l*mu_l, upwind alpha_l*F, temporal,
                 Schiller-Naumann implicit drag on the diagonal, Dirichlet
                 0 walls); the drag coefficient is computed on the
                 device (kcfd_drag_implicit).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from .mesh import (build_fvmesh_from_arrays, structured_hex_arrays, structured_quad_arrays,
                   tet_arrays)

# two-fluid constants (models/two_fluid.py:36-47)
RHO_L, RHO_G, MU_L, MU_G, D_B = 998.2, 1.225, 1.003e-3, 1.789e-5, 0.005


@dataclasses.dataclass
class Case:
    name: str
    mesh: object
    recipe: list          # [(op, kwargs)]
    precond: str
    rhs: np.ndarray | None = None   # final rhs override (set after assembly)
    meta: dict = dataclasses.field(default_factory=dict)


# --------------------------------------------------------------------- helpers

def _face_sign_2d(mesh):
    """+1 where the stored normal equals (dy,-dx)/L of the face's node order."""
    fn = mesh.face_nodes
    p0 = mesh.nodes[fn[:, 0]]
    p1 = mesh.nodes[fn[:, 1]]
    raw = np.stack([p1[:, 1] - p0[:, 1], -(p1[:, 0] - p0[:, 0])], 1)
    return np.where(np.sum(raw * mesh.face_normal[:, :2], axis=1) >= 0, 1.0, -1.0)


def _stream_function(rng, nmodes=8):
    a = rng.normal(0.0, 1.0, nmodes)
    m = rng.integers(1, 5, nmodes)
    k = rng.integers(1, 5, nmodes)

    def psi(x, y):
        return np.sum(a[None, :] * np.sin(np.pi * m[None, :] * x[:, None]) *
                      np.sin(np.pi * k[None, :] * y[:, None]), axis=1)

    def vel(x, y):
        # u = (d psi/dy, -d psi/dx)
        sx = np.sin(np.pi * m[None, :] * x[:, None])
        cx = np.cos(np.pi * m[None, :] * x[:, None])
        sy = np.sin(np.pi * k[None, :] * y[:, None])
        cy = np.cos(np.pi * k[None, :] * y[:, None])
        u = np.sum(a * sx * np.pi * k * cy, axis=1)
        v = -np.sum(a * np.pi * m * cx * sy, axis=1)
        return u, v

    return psi, vel


def _vector_potential(rng, nmodes=8):
    c = rng.normal(0.0, 1.0, (nmodes, 3))
    w = rng.integers(1, 4, (nmodes, 3))
    ph = rng.uniform(0, 2 * np.pi, nmodes)

    def A(p):
        arg = np.pi * (p @ w.T) + ph[None, :]          # (np, nmodes)
        return np.sin(arg) @ c                          # (np, 3)

    return A


def _circulation_flux(mesh, A):
    """Face flux = circulation of A around the face (trapezoid per edge), oriented
    to the stored normal.  Each edge integral is antisymmetric, so the flux
    out of every cell sums to zero up to rounding."""
    fn_all = mesh.face_nodes
    P = mesh.nodes
    Av = A(P)
    out = np.zeros(len(fn_all))
    nnodes = (fn_all >= 0).sum(axis=1)      # mixed meshes pad triangles with -1
    for k in np.unique(nnodes):
        sel = np.nonzero(nnodes == k)[0]
        fn = fn_all[sel, :k]
        circ = np.zeros(len(fn))
        for j in range(k):
            a = fn[:, j]
            b = fn[:, (j + 1) % k]
            circ += np.sum(0.5 * (Av[a] + Av[b]) * (P[b] - P[a]), axis=1)
        # Newell normal of the node order (unnormalised) vs the stored normal
        pts = P[fn]
        nxt = np.roll(pts, -1, axis=1)
        nn = np.stack([np.sum((pts[:, :, 1] - nxt[:, :, 1]) * (pts[:, :, 2] + nxt[:, :, 2]), 1),
                       np.sum((pts[:, :, 2] - nxt[:, :, 2]) * (pts[:, :, 0] + nxt[:, :, 0]), 1),
                       np.sum((pts[:, :, 0] - nxt[:, :, 0]) * (pts[:, :, 1] + nxt[:, :, 1]), 1)], 1)
        sgn = np.where(np.sum(nn * mesh.face_normal[sel], axis=1) >= 0, 1.0, -1.0)
        out[sel] = sgn * circ
    return out


def dirichlet_all(mesh, value=0.0):
    bc = {nm: {"type": "dirichlet"} for nm in mesh.boundary_patches}
    phib = {nm: np.full(len(f), value) for nm, f in mesh.boundary_patches.items()}
    return bc, phib


# ------------------------------------------------------------------ configs

def quad_mesh(nx, ny=None):
    ny = nx if ny is None else ny
    nodes, cells, bf = structured_quad_arrays(0.0, 0.0, 1.0, 1.0, nx, ny)
    return build_fvmesh_from_arrays(nodes, cells, bf, ndim=2)


def hex_mesh(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    nodes, cells, faces, bf = structured_hex_arrays(1.0, 1.0, 1.0, nx, ny, nz)
    return build_fvmesh_from_arrays(nodes, cells, bf, ndim=3, cell_face_node_list=faces)


def quad_convdiff(n=128, seed=1001, mesh=None) -> Case:
    rng = np.random.default_rng(seed)
    mesh = quad_mesh(n) if mesh is None else mesh
    nc = mesh.n_cells
    gamma = np.exp(rng.normal(math.log(1e-2), 0.5, nc))
    psi, vel = _stream_function(rng)
    fn = mesh.face_nodes
    P = mesh.nodes
    F = _face_sign_2d(mesh) * (psi(P[fn[:, 1], 0], P[fn[:, 1], 1]) - psi(P[fn[:, 0], 0], P[fn[:, 0], 1]))
    u, v = vel(mesh.cell_center[:, 0], mesh.cell_center[:, 1])
    F = F / np.max(np.sqrt(u * u + v * v))
    b = rng.uniform(-1.0, 1.0, nc)
    bc, phib = dirichlet_all(mesh)
    recipe = [("diffusion", dict(gamma=gamma)), ("upwind", dict(F=F)),
              ("boundary", dict(gamma=gamma, F=F, phib=phib, bc=bc))]
    return Case("quad_convdiff", mesh, recipe, "jacobi", rhs=b,
                meta=dict(n=n, seed=seed))


def hex_pressure(n=128, seed=1002, mesh=None) -> Case:
    rng = np.random.default_rng(seed)
    mesh = hex_mesh(n) if mesh is None else mesh
    nc = mesh.n_cells
    aP = np.exp(rng.normal(0.0, 1.0, nc))
    dP = mesh.cell_volume / aP
    b = rng.uniform(-1.0, 1.0, nc)
    pdir = [nm for nm in mesh.boundary_patches if nm == "outlet"]
    recipe = [("pressure", dict(dP=dP, rho=1.0, dirichlet=pdir))]
    return Case("hex_pressure", mesh, recipe, "ilu0", rhs=b, meta=dict(n=n, seed=seed))


def pressure_case(mesh, seed, dirichlet=None) -> Case:
    """SIMPLE pressure correction on any mesh (the hex_pressure recipe):
    a_P ~ lognormal(0, 1), d_P = V / a_P, p' = 0 on ``dirichlet`` patches
    (default: 'outlet' if present, else the right / first patch),
    b ~ U[-1, 1]."""
    rng = np.random.default_rng(seed)
    nc = mesh.n_cells
    aP = np.exp(rng.normal(0.0, 1.0, nc))
    dP = mesh.cell_volume / aP
    b = rng.uniform(-1.0, 1.0, nc)
    if dirichlet is None:
        names = list(mesh.boundary_patches)
        dirichlet = [nm for nm in ("outlet", "right") if nm in names][:1] or names[:1]
    recipe = [("pressure", dict(dP=dP, rho=1.0, dirichlet=list(dirichlet)))]
    return Case(f"{'hex' if mesh.ndim == 3 else 'quad'}_pressure", mesh, recipe, "ilu0", rhs=b,
                meta=dict(seed=seed))


def hex_momentum(n=128, seed=1002, mesh=None) -> Case:
    rng = np.random.default_rng(seed + 7)
    mesh = hex_mesh(n) if mesh is None else mesh
    nc = mesh.n_cells
    gamma = np.full(nc, 1e-2)
    F = _circulation_flux(mesh, _vector_potential(rng))
    F = F / np.max(np.abs(F) / mesh.face_area)
    x = mesh.cell_center
    phi = np.sin(2 * np.pi * x[:, 0]) * np.cos(np.pi * x[:, 1]) + 0.5 * np.sin(3 * np.pi * x[:, 2])
    bc, phib = dirichlet_all(mesh)
    src = rng.uniform(-1.0, 1.0, nc)
    recipe = [("diffusion", dict(gamma=gamma)), ("upwind", dict(F=F)),
              ("boundary", dict(gamma=gamma, F=F, phib=phib, bc=bc)),
              ("source", dict(S=src)),
              ("muscl", dict(phi=phi, phib=phib, F=F, limiter="van_leer"))]
    return Case("hex_momentum", mesh, recipe, "ilu0", meta=dict(n=n, seed=seed))


def hybrid_mesh(nx=64, ny=None, nz=None, tet_fraction=0.5, L=(2.0, 0.1, 0.1)):
    """Mixed hex/tet duct of mesh/hybrid_mesh_generator.py:93-270 (the
    geometry of verification/case22_hybrid_mesh.py): hexes on the inlet
    half, 12 tets per hex on the outlet half."""
    from .mesh import generate_hybrid_hex_tet_mesh
    ny = max(1, nx // 2) if ny is None else ny
    nz = ny if nz is None else nz
    return generate_hybrid_hex_tet_mesh(L[0], L[1], L[2], nx, ny, nz, tet_fraction)


def hybrid_pressure(nx=64, ny=None, nz=None, seed=1007, mesh=None) -> Case:
    """The hex_pressure recipe on the mixed hex/tet duct (p' = 0 at the
    outlet)."""
    mesh = hybrid_mesh(nx, ny, nz) if mesh is None else mesh
    case = pressure_case(mesh, seed, dirichlet=["outlet"])
    case.name = "hybrid_pressure"
    return case


def hybrid_momentum(nx=64, ny=None, nz=None, seed=1008, mesh=None) -> Case:
    """The hex_momentum recipe on the mixed hex/tet duct: diffusion, upwind
    convection by a divergence-free circulation flux through the hex, tet
    and interface faces, Dirichlet 0, source, van Leer MUSCL correction."""
    mesh = hybrid_mesh(nx, ny, nz) if mesh is None else mesh
    case = hex_momentum(seed=seed, mesh=mesh)
    case.name = "hybrid_momentum"
    return case


def tet_mesh(npts_target=200_000, seed=1003, jitter=0.3):
    """Delaunay tetrahedra of a jittered lattice; cells randomly permuted,
    then RCM-reordered."""
    from scipy.spatial import Delaunay
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    rng = np.random.default_rng(seed)
    m = max(2, int(round(npts_target ** (1.0 / 3.0))))
    nx, ny, nz = m, m, max(2, int(round(npts_target / (m * m))))
    h = np.array([1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1)])
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)
    pts = idx * h
    jit = rng.uniform(-jitter, jitter, pts.shape) * h
    onb = (idx == 0) | (idx == np.array([nx - 1, ny - 1, nz - 1]))
    jit[onb] = 0.0            # keep boundary points on the boundary planes
    pts = pts + jit
    tri = Delaunay(pts)
    tets = tri.simplices.astype(np.int64)
    # drop slivers of (near) zero volume
    p = pts[tets]
    vol6 = np.abs(np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0]))
    tets = tets[vol6 > 1e-12 * np.prod(h)]
    perm = rng.permutation(len(tets))
    tets = tets[perm]
    mesh0 = build_fvmesh_from_arrays(pts, tets, {}, ndim=3, cell_face_node_list=tet_arrays(tets))
    o = mesh0.owner.astype(np.int64)
    nb = mesh0.neighbour.astype(np.int64)
    it = nb >= 0
    n = mesh0.n_cells
    G = csr_matrix((np.ones(2 * it.sum()), (np.r_[o[it], nb[it]], np.r_[nb[it], o[it]])), shape=(n, n))
    order = reverse_cuthill_mckee(G, symmetric_mode=True)
    tets = tets[order]
    return build_fvmesh_from_arrays(pts, tets, {}, ndim=3, cell_face_node_list=tet_arrays(tets))


def tet_twofluid(npts=200_000, seed=1003, mesh=None) -> Case:
    mesh = tet_mesh(npts, seed) if mesh is None else mesh
    rng = np.random.default_rng(seed + 11)
    nc = mesh.n_cells
    nd = 3
    alpha_g = rng.uniform(0.05, 0.6, nc)
    alpha_l = 1.0 - alpha_g
    U_l = rng.normal(0.0, 0.2, (nc, nd))
    U_g = U_l + rng.normal(0.0, 0.2, (nc, nd))
    Ub = {nm: np.zeros((len(f), nd)) for nm, f in mesh.boundary_patches.items()}
    dt = 1e-3
    gamma = alpha_l * MU_L
    phi_old = U_l[:, 0].copy()
    bc, phib = dirichlet_all(mesh)
    recipe = [("mass_flux", dict(U=U_l, Ub=Ub, rho=RHO_L)),
              ("diffusion", dict(gamma=gamma)),
              ("upwind_scaled", dict(scale_owner=alpha_l)),
              ("cell_terms", dict(kd=alpha_l * RHO_L / dt, ks=phi_old)),
              # Schiller-Naumann implicit drag (models/closure.py:72-92, applied
              # at models/two_fluid.py:592-596): K computed on the device
              ("drag", dict(alpha_g=alpha_g, U_g=U_g, U_l=U_l, comp=0, rho_l=RHO_L, d_b=D_B, mu_l=MU_L)),
              ("boundary_scaled", dict(gamma=gamma, scale_owner=alpha_l, phib=phib, bc=bc))]
    return Case("tet_twofluid", mesh, recipe, "jacobi",
                meta=dict(npts=npts, seed=seed, alpha_g=alpha_g))


# --------------------------------------------------------------- batch mix

def batch_specs(nsys=512, seed=4000, nmin=2_000, nmax=250_000):
    """Sizes log-uniform in [nmin, nmax]; types 50% quad, 30% hex, 20% tet."""
    out = []
    for i in range(nsys):
        rng = np.random.default_rng(seed + i)
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        u = rng.uniform()
        kind = "quad" if u < 0.5 else ("hex" if u < 0.8 else "tet")
        out.append((kind, n, seed + i))
    return out


def case_of_size(kind, ncells, seed):
    if kind == "quad":
        s = max(4, int(round(math.sqrt(ncells))))
        return quad_convdiff(s, seed)
    if kind == "hex":
        s = max(3, int(round(ncells ** (1.0 / 3.0))))
        return hex_pressure(s, seed)
    npts = max(64, int(round(ncells / 5.3)))
    return tet_twofluid(npts, seed)


# ---------------------------------------------------------- device assembly

class DeviceRecipe:
    """A case's assembly recipe with its arrays uploaded once; ``assemble``
    plays it into a matrix of the mesh's structure (many systems on one
    mesh share the uploads)."""

    def __init__(self, dmesh, case: Case):
        from . import _lib as L
        from ._lib import DevBuf, check, lib

        self.dmesh = dmesh
        self.case = case
        ctx = dmesh.ctx
        mesh = case.mesh
        self.keep = []
        self.ops = []

        def dev(a):
            b = DevBuf.from_array(ctx, np.ascontiguousarray(a, np.float64))
            self.keep.append(b)
            return b.p

        F_dev = None
        F_scaled = None
        for op, kw in case.recipe:
            if op == "mass_flux":
                F_dev = DevBuf(ctx, mesh.n_faces)
                self.keep.append(F_dev)
                check(lib().kcfd_mass_flux(dmesh.h, dev(kw["U"]),
                                           dev(dmesh.boundary_array(kw["Ub"], width=mesh.ndim)),
                                           float(kw["rho"]), F_dev.p))
                ctx.sync()
            elif op == "diffusion":
                g = dev(kw["gamma"])
                self.ops.append(lambda A, g=g: check(lib().kcfd_asm_diffusion(dmesh.h, A.h, g)))
            elif op == "upwind":
                f = dev(kw["F"])
                self.ops.append(lambda A, f=f: check(lib().kcfd_asm_upwind(dmesh.h, A.h, f)))
            elif op == "upwind_scaled":
                F_scaled = F_dev.download() * np.asarray(kw["scale_owner"])[mesh.owner]
                f = dev(F_scaled)
                self.ops.append(lambda A, f=f: check(lib().kcfd_asm_upwind(dmesh.h, A.h, f)))
            elif op in ("boundary", "boundary_scaled"):
                F = kw["F"] if op == "boundary" else F_scaled
                kinds = np.array([{"dirichlet": L.BC_DIRICHLET, "neumann": L.BC_NEUMANN,
                                   "zero_gradient": L.BC_ZERO_GRADIENT}[
                                      kw["bc"].get(nm, {"type": "zero_gradient"})["type"]]
                                  for nm in dmesh.patch_names], np.int32)
                vals = np.array([float(kw["bc"].get(nm, {}).get("value", 0.0)) for nm in dmesh.patch_names])
                args = (dev(kw["gamma"]), dev(F), dev(dmesh.boundary_array(kw["phib"])))
                self.keep += [kinds, vals]
                self.ops.append(lambda A, a=args, k=kinds, v=vals: check(lib().kcfd_asm_boundary(
                    dmesh.h, A.h, a[0], a[1], a[2], k.ctypes.data, v.ctypes.data)))
            elif op == "pressure":
                pd = np.array([1 if nm in kw["dirichlet"] else 0 for nm in dmesh.patch_names], np.int32)
                dp = dev(kw["dP"])
                self.keep.append(pd)
                self.ops.append(lambda A, dp=dp, r=float(kw["rho"]), pd=pd: check(
                    lib().kcfd_asm_pressure(dmesh.h, A.h, dp, r, pd.ctypes.data)))
            elif op == "source":
                src = dev(kw["S"])
                self.ops.append(lambda A, src=src: check(lib().kcfd_asm_source(dmesh.h, A.h, src)))
            elif op == "drag":
                K = DevBuf(ctx, mesh.n_cells)
                self.keep.append(K)
                U_g = np.ascontiguousarray(kw["U_g"], np.float64)
                check(lib().kcfd_drag_implicit(ctx.h, mesh.n_cells, U_g.shape[1], dev(kw["alpha_g"]),
                                               float(kw["rho_l"]), dev(U_g), dev(kw["U_l"]), float(kw["d_b"]),
                                               float(kw["mu_l"]), K.p))
                ks = dev(U_g[:, kw["comp"]])
                self.ops.append(lambda A, kd=K.p, ks=ks: check(lib().kcfd_asm_cell_terms(dmesh.h, A.h, kd, ks)))
            elif op == "cell_terms":
                kd, ks = dev(kw["kd"]), dev(kw["ks"])
                self.ops.append(lambda A, kd=kd, ks=ks: check(lib().kcfd_asm_cell_terms(dmesh.h, A.h, kd, ks)))
            elif op == "temporal":
                po = dev(kw["phi_old"])
                self.ops.append(lambda A, po=po, r=float(kw["rho"]), dt=float(kw["dt"]): check(
                    lib().kcfd_asm_temporal(dmesh.h, A.h, r, dt, po)))
            elif op == "muscl":
                phib = dev(dmesh.boundary_array(kw["phib"]))
                g = DevBuf(ctx, mesh.n_cells * mesh.ndim)
                c = DevBuf(ctx, mesh.n_cells)
                self.keep += [g, c]
                phi, F = dev(kw["phi"]), dev(kw["F"])
                lim = L.LIMITERS[kw["limiter"]]

                def muscl(A, phi=phi, phib=phib, F=F, g=g, c=c, lim=lim):
                    check(lib().kcfd_grad_gg(dmesh.h, phi, phib, g.p))
                    check(lib().kcfd_muscl(dmesh.h, phi, F, g.p, lim, c.p))
                    A.set_rhs(A.get_rhs() + c.download())
                self.ops.append(muscl)
            else:
                raise ValueError(f"unknown recipe op {op}")

    def assemble(self, A, rhs=None):
        """Play the recipe into A (reset first); rhs overrides the assembled
        right-hand side (default: the case's own override, if any)."""
        from ._lib import check, lib
        check(lib().kcfd_asm_reset(A.h))
        for f in self.ops:
            f(A)
        self.dmesh.ctx.sync()
        rhs = self.case.rhs if rhs is None else rhs
        if rhs is not None:
            A.set_rhs(rhs)
        return A

    def close(self):
        for b in self.keep:
            if hasattr(b, "free"):
                b.free()
        self.keep = []


def assemble_device(dmesh, case: Case, ctx=None):
    """Play the recipe on the device; returns the DeviceMatrix (rhs set)."""
    rec = DeviceRecipe(dmesh, case)
    try:
        return rec.assemble(dmesh.new_system())
    finally:
        rec.close()
