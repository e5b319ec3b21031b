"""Cell-centred fields with per-patch boundary values.

Same attributes and semantics as the reference's ScalarField / VectorField
(core/fields.py:12-155): ``values`` (n,) or (n, ndim), ``boundary_values``
{patch: array}, ``old_values``; the mesh is the SoA ``kcfd_amd.mesh.FVMesh``.
"""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np


class ScalarField:
    """core/fields.py:12-74."""

    def __init__(self, mesh, name: str = "scalar", default: float = 0.0):
        self.mesh = mesh
        self.name = name
        self.values = np.full(mesh.n_cells, default, dtype=np.float64)
        self.boundary_values: Dict[str, np.ndarray] = {
            nm: np.full(len(f), default, dtype=np.float64) for nm, f in mesh.boundary_patches.items()}
        self.old_values: Optional[np.ndarray] = None

    def copy(self) -> "ScalarField":
        sf = ScalarField(self.mesh, self.name)
        sf.values = self.values.copy()
        for nm in self.boundary_values:
            sf.boundary_values[nm] = self.boundary_values[nm].copy()
        if self.old_values is not None:
            sf.old_values = self.old_values.copy()
        return sf

    def store_old(self):
        self.old_values = self.values.copy()

    def set_uniform(self, value: float):
        self.values[:] = value
        for nm in self.boundary_values:
            self.boundary_values[nm][:] = value

    def set_boundary(self, patch_name: str, value):
        if patch_name not in self.boundary_values:
            return
        if np.isscalar(value):
            self.boundary_values[patch_name][:] = value
        else:
            self.boundary_values[patch_name] = np.array(value, dtype=np.float64)

    def get_face_value(self, face_idx: int) -> float:
        if self.mesh.neighbour[face_idx] == -1:
            for nm, fids in self.mesh.boundary_patches.items():
                hit = np.nonzero(np.asarray(fids) == face_idx)[0]
                if hit.size:
                    return self.boundary_values[nm][hit[0]]
        return self.values[self.mesh.owner[face_idx]]

    def max(self) -> float:
        return float(np.max(self.values))

    def min(self) -> float:
        return float(np.min(self.values))

    def mean(self) -> float:
        return float(np.mean(self.values))


class VectorField:
    """core/fields.py:77-155 (ndim components per cell)."""

    def __init__(self, mesh, name: str = "vector", default: np.ndarray = None):
        self.mesh = mesh
        self.name = name
        nd = getattr(mesh, "ndim", 2)
        default = np.zeros(nd) if default is None else np.asarray(default, dtype=np.float64)
        if default.shape[0] < nd:            # a short default is zero-padded to ndim
            default = np.concatenate([default, np.zeros(nd - default.shape[0])])
        self.values = np.tile(default, (mesh.n_cells, 1))
        self.boundary_values: Dict[str, np.ndarray] = {
            nm: np.tile(default, (len(f), 1)) for nm, f in mesh.boundary_patches.items()}
        self.old_values: Optional[np.ndarray] = None

    # component views: x, y always; z reads zeros and ignores writes below 3D
    @property
    def x(self) -> np.ndarray:
        return self.values[:, 0]

    @x.setter
    def x(self, val):
        self.values[:, 0] = val

    @property
    def y(self) -> np.ndarray:
        return self.values[:, 1]

    @y.setter
    def y(self, val):
        self.values[:, 1] = val

    @property
    def z(self) -> np.ndarray:
        if self.values.shape[1] < 3:
            return np.zeros(self.values.shape[0])
        return self.values[:, 2]

    @z.setter
    def z(self, val):
        if self.values.shape[1] >= 3:
            self.values[:, 2] = val

    def copy(self) -> "VectorField":
        vf = VectorField(self.mesh, self.name)
        vf.values = self.values.copy()
        for nm in self.boundary_values:
            vf.boundary_values[nm] = self.boundary_values[nm].copy()
        if self.old_values is not None:
            vf.old_values = self.old_values.copy()
        return vf

    def store_old(self):
        self.old_values = self.values.copy()

    def set_uniform(self, value):
        v = np.asarray(value, dtype=np.float64)
        self.values[:] = v
        for nm in self.boundary_values:
            self.boundary_values[nm][:] = v

    def set_boundary(self, patch_name: str, value):
        if patch_name not in self.boundary_values:
            return
        v = np.asarray(value, dtype=np.float64)
        if v.ndim == 1:
            self.boundary_values[patch_name][:] = v
        else:
            self.boundary_values[patch_name] = v.copy()

    def magnitude(self) -> np.ndarray:
        return np.sqrt(np.sum(self.values ** 2, axis=1))
