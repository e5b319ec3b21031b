"""Per-mesh device state shared by the drop-in operators."""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from ..device import DeviceMesh


def device_mesh(mesh) -> DeviceMesh:
    """The DeviceMesh of ``mesh`` (built once, cached on the mesh object)."""
    dm = getattr(mesh, "_kcfd_device_mesh", None)
    if dm is None:
        dm = DeviceMesh(mesh, L.default_context())
        mesh._kcfd_device_mesh = dm
    return dm


class Scratch:
    """Short-lived device copies of host inputs for one operator call."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.bufs = []

    def put(self, a, dtype=np.float64):
        b = L.DevBuf.from_array(self.ctx, np.ascontiguousarray(a, dtype=dtype))
        self.bufs.append(b)
        return b

    def empty(self, n, dtype=np.float64):
        b = L.DevBuf(self.ctx, n, dtype)
        self.bufs.append(b)
        return b

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.ctx.sync()
        for b in self.bufs:
            b.free()
        self.bufs.clear()


def values_of(x):
    return x.values if hasattr(x, "values") else np.asarray(x, dtype=np.float64)
