"""Per-mesh device state shared by the drop-in operators."""

from __future__ import annotations

import threading

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


class _Slot:
    """A pooled device buffer seen as an array of ``n`` elements."""

    def __init__(self, buf, n, dtype):
        self.buf = buf
        self.p = buf.p
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize

    def download(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            L.check(L.lib().kcfd_buf_download(self.buf.ctx.h, L.ptr(out), self.p, self.nbytes))
        return out


_POOL: dict = {}          # id(ctx) -> free device buffers
_POOL_KEEP = 16           # buffers kept per context between calls
_POOL_LOCK = threading.Lock()


class Scratch:
    """Device copies of host inputs for one operator call, in buffers reused
    from call to call (a per-context pool: no device allocation once the
    pool holds large enough buffers).  Everything runs on the context's one
    stream, so a buffer handed back while a kernel still reads it is only
    overwritten by a later upload on that stream, after the kernel."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.bufs = []

    def _take(self, nbytes: int):
        with _POOL_LOCK:
            pool = _POOL.setdefault(id(self.ctx), [])
            fits = [i for i, b in enumerate(pool) if b.nbytes >= nbytes]
            buf = pool.pop(min(fits, key=lambda i: pool[i].nbytes)) if fits else None
        if buf is None:
            buf = L.DevBuf(self.ctx, max(int(nbytes), 8), np.uint8)
        self.bufs.append(buf)
        return buf

    def put(self, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype).ravel()
        buf = self._take(a.nbytes)
        if a.nbytes:
            L.check(L.lib().kcfd_buf_upload(self.ctx.h, buf.p, L.ptr(a), a.nbytes))
        return _Slot(buf, a.size, dtype)

    def empty(self, n, dtype=np.float64):
        return _Slot(self._take(int(n) * np.dtype(dtype).itemsize), n, dtype)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        with _POOL_LOCK:
            pool = _POOL.setdefault(id(self.ctx), [])
            pool.extend(self.bufs)
            pool.sort(key=lambda b: -b.nbytes)
            drop = pool[_POOL_KEEP:]
            del pool[_POOL_KEEP:]
        for b in drop:
            b.free()
        self.bufs.clear()


def values_of(x):
    return x.values if hasattr(x, "values") else np.asarray(x, dtype=np.float64)
