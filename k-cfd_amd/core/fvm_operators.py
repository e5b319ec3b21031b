"""Drop-in FVM operators over the device-resident CSR system.

Mirrors core/fvm_operators.py (names, arguments, order of effects):

* ``FVMSystem(n)`` keeps the reference's public surface -- ``n``, ``rows``,
  ``cols``, ``vals``, ``rhs``, ``diag``, ``add_diagonal``,
  ``add_off_diagonal``, ``add_source``, ``to_sparse``, ``reset``
  (fvm_operators.py:13-56).  The operators below do not append COO triplets:
  they assemble on the device, straight into the canonical CSR of the mesh
  (face-parallel scatter through the static face->slot map plus a cell
  gather for diagonal, diagonal tracker and rhs).  Entries added by hand
  with ``add_*`` are kept in the host COO lists exactly as the reference
  does and merged when the matrix is materialised.
* ``rhs`` and ``diag`` are numpy arrays the caller may mutate in place
  (models/two_fluid.py:606-609 floors the tracker, case6_muscl.py:117 adds to
  rhs).  Reading them copies the device arrays to the host; the next device
  operator uploads them back first, so either side can be edited.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse

from .. import _lib as L
from ..device import DeviceMatrix
from ._dev import Scratch, device_mesh, values_of

_BC = {"dirichlet": L.BC_DIRICHLET, "neumann": L.BC_NEUMANN, "zero_gradient": L.BC_ZERO_GRADIENT}


class FVMSystem:
    """fvm_operators.py:13-56 with a device-resident matrix."""

    def __init__(self, n: int):
        self.n = n
        self.rows: list = []
        self.cols: list = []
        self.vals: list = []
        self._mesh = None
        self._dm = None
        self._A: DeviceMatrix | None = None
        self._host = {"rhs": np.zeros(n, dtype=np.float64), "diag": np.zeros(n, dtype=np.float64)}
        self._owner = {"rhs": "host", "diag": "host"}   # which copy is authoritative

    # -- device binding ---------------------------------------------------
    def _bind(self, mesh) -> DeviceMatrix:
        if mesh.n_cells != self.n:
            raise ValueError(f"mesh has {mesh.n_cells} cells, system has n={self.n}")
        if self._A is None:
            self._mesh = mesh
            self._dm = device_mesh(mesh)
            self._A = self._dm.new_system()
            # whatever the host arrays hold becomes the device starting point
            self._push("rhs")
            self._push("diag")
        elif mesh is not self._mesh and getattr(mesh, "_kcfd_device_mesh", None) is not self._dm:
            raise ValueError("FVMSystem is bound to another mesh")
        self._sync_to_device()
        return self._A

    def _push(self, key):
        a = np.ascontiguousarray(self._host[key], dtype=np.float64)
        if key == "rhs":
            self._A.set_rhs(a)
        else:
            self._A.set_diag_tracker(a)
        self._owner[key] = "dev"

    def _sync_to_device(self):
        for key in ("rhs", "diag"):
            if self._owner[key] == "host":
                self._push(key)

    def _pull(self, key):
        if self._A is not None and self._owner[key] == "dev":
            self._host[key] = self._A.get_rhs() if key == "rhs" else self._A.get_diag_tracker()
            self._owner[key] = "host"
        return self._host[key]

    @property
    def rhs(self) -> np.ndarray:
        return self._pull("rhs")

    @rhs.setter
    def rhs(self, v):
        self._host["rhs"] = np.asarray(v, dtype=np.float64)
        self._owner["rhs"] = "host"

    @property
    def diag(self) -> np.ndarray:
        return self._pull("diag")

    @diag.setter
    def diag(self, v):
        self._host["diag"] = np.asarray(v, dtype=np.float64)
        self._owner["diag"] = "host"

    @property
    def device_matrix(self) -> DeviceMatrix | None:
        if self._A is not None:
            self._sync_to_device()
        return self._A

    # -- reference COO surface ---------------------------------------------
    def add_diagonal(self, i: int, val: float):
        self.rows.append(i)
        self.cols.append(i)
        self.vals.append(val)
        self.diag[i] += val

    def add_off_diagonal(self, i: int, j: int, val: float):
        self.rows.append(i)
        self.cols.append(j)
        self.vals.append(val)

    def add_source(self, i: int, val: float):
        self.rhs[i] += val

    def has_host_entries(self) -> bool:
        return len(self.vals) > 0

    def to_sparse(self) -> sparse.csr_matrix:
        """Canonical CSR: device-assembled values (if any) plus the host COO
        triplets, duplicates summed (fvm_operators.py:43-48)."""
        if self._A is None:
            return sparse.coo_matrix((self.vals, (self.rows, self.cols)), shape=(self.n, self.n)).tocsr()
        A = self._A.to_scipy()
        if self.vals:
            B = sparse.coo_matrix((self.vals, (self.rows, self.cols)), shape=(self.n, self.n)).tocsr()
            A = (A + B).tocsr()
            A.sum_duplicates()
        return A

    def reset(self):
        self.rows.clear()
        self.cols.clear()
        self.vals.clear()
        self._host["rhs"] = np.zeros(self.n)
        self._host["diag"] = np.zeros(self.n)
        self._owner = {"rhs": "host", "diag": "host"}
        if self._A is not None:
            L.check(L.lib().kcfd_asm_reset(self._A.h))


def _ctx():
    return L.default_context()


def diffusion_operator(mesh, gamma, system: FVMSystem):
    """fvm_operators.py:59-100 on the device."""
    A = system._bind(mesh)
    dm = system._dm
    with Scratch(dm.ctx) as s:
        g = s.put(values_of(gamma))
        L.check(L.lib().kcfd_asm_diffusion(dm.h, A.h, g.p))


def convection_operator_upwind(mesh, mass_flux, system: FVMSystem):
    """fvm_operators.py:103-129 on the device."""
    A = system._bind(mesh)
    dm = system._dm
    with Scratch(dm.ctx) as s:
        F = s.put(mass_flux)
        L.check(L.lib().kcfd_asm_upwind(dm.h, A.h, F.p))


def temporal_operator(mesh, rho: float, dt: float, phi_old, system: FVMSystem):
    """fvm_operators.py:132-148 on the device."""
    A = system._bind(mesh)
    dm = system._dm
    with Scratch(dm.ctx) as s:
        p = s.put(phi_old)
        L.check(L.lib().kcfd_asm_temporal(dm.h, A.h, float(rho), float(dt), p.p))


def source_term(mesh, source_values, system: FVMSystem):
    """fvm_operators.py:151-161 on the device."""
    A = system._bind(mesh)
    dm = system._dm
    with Scratch(dm.ctx) as s:
        p = s.put(source_values)
        L.check(L.lib().kcfd_asm_source(dm.h, A.h, p.p))


def linearized_source(mesh, Sp, Su, system: FVMSystem):
    """fvm_operators.py:164-175 on the device."""
    A = system._bind(mesh)
    dm = system._dm
    with Scratch(dm.ctx) as s:
        a, b = s.put(Sp), s.put(Su)
        L.check(L.lib().kcfd_asm_linearized(dm.h, A.h, a.p, b.p))


def apply_boundary_conditions(mesh, phi, gamma, mass_flux, system: FVMSystem, bc_types: dict):
    """fvm_operators.py:178-225 on the device.  Patches missing from
    ``bc_types`` are zero_gradient; an unknown type string contributes
    nothing, as in the reference's if/elif chain."""
    A = system._bind(mesh)
    dm = system._dm
    kinds = np.array([_BC.get(bc_types.get(nm, {"type": "zero_gradient"})["type"], L.BC_NONE)
                      for nm in dm.patch_names], dtype=np.int32)
    vals = np.array([float(bc_types.get(nm, {}).get("value", 0.0)) for nm in dm.patch_names])
    bvals = getattr(phi, "boundary_values", {}) or {}
    phib = dm.boundary_array(bvals)
    with Scratch(dm.ctx) as s:
        g, F, pb = s.put(values_of(gamma)), s.put(mass_flux), s.put(phib)
        L.check(L.lib().kcfd_asm_boundary(dm.h, A.h, g.p, F.p, pb.p, kinds.ctypes.data,
                                          vals.ctypes.data))


def under_relax(system: FVMSystem, phi, alpha: float):
    """fvm_operators.py:228-240 (reads the diagonal tracker) on the device."""
    if alpha >= 1.0:
        return
    if system._A is None:
        # nothing assembled on the device yet: the tracker lives on the host
        # the reference's per-cell add_diagonal / add_source, as arrays
        ap = system.diag.copy()
        k = ap * (1.0 - alpha) / alpha
        idx = range(system.n)
        system.rows.extend(idx)
        system.cols.extend(idx)
        system.vals.extend(k.tolist())
        system.diag[:] = ap + k
        system.rhs[:] += ap * (1.0 - alpha) / alpha * values_of(phi)
        return
    system._sync_to_device()
    A = system._A
    with Scratch(system._dm.ctx) as s:
        p = s.put(values_of(phi))
        L.check(L.lib().kcfd_asm_under_relax(A.h, p.p, float(alpha)))
