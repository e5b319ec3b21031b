"""Device-resident objects over the C-ABI: matrices, meshes, solves, batches.

Everything here drives libkcfd.so; numpy is used only to prepare inputs
(mesh structure, slot maps) and to hold results copied back.
"""

from __future__ import annotations

import ctypes

import numpy as np
from scipy import sparse

from . import _lib as L
from ._lib import DevBuf, check, lib, ptr


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def fp64_continuation(mat, b_ptr, x_ptr, opts, mixed_res):
    """Continue a mixed solve that did not converge in FP64 from x (the
    best iterate).  The matrix set up for MIXED also holds its FP64 values
    and preconditioner, so no set-up is repeated."""
    o = L.SolveOpts()
    ctypes.pointer(o)[0] = opts
    o.precision = L.FP64
    r = L.SolveResult()
    check(lib().kcfd_solve(mat.h, b_ptr, x_ptr, ctypes.byref(o), ctypes.byref(r)))
    out = L.SolveResult()
    out.info = r.info
    out.iters = mixed_res.iters + r.iters
    out.half_step = r.half_step
    out.refinements = mixed_res.refinements
    out.rel_resid = r.rel_resid
    out.bytes = mixed_res.bytes + r.bytes
    out.ms = mixed_res.ms + r.ms
    out.fallback = 1
    return out


def default_opts(rtol=1e-8, atol=0.0, maxiter=0, precond="jacobi", precision="fp64",
                 inner_rtol=0.1, max_refine=1000, check_every=8, grid_per_cu=0,
                 trsv_per_cu=0, tile_walk=0, method="bicgstab", restart=0, rows_per_thread=0,
                 fold_colour0=0) -> L.SolveOpts:
    o = L.SolveOpts()
    o.rtol = float(rtol)
    o.atol = float(atol)
    o.maxiter = int(maxiter or 0)
    o.precond = L.PRECONDS[precond] if isinstance(precond, str) else int(precond)
    o.precision = {"fp64": L.FP64, "mixed": L.MIXED}[precision] if isinstance(precision, str) \
        else int(precision)
    o.inner_rtol = float(inner_rtol)
    o.max_refine = int(max_refine)
    o.check_every = int(check_every)
    o.grid_per_cu = int(grid_per_cu)
    o.trsv_per_cu = int(trsv_per_cu)
    o.tile_walk = int(tile_walk)
    o.method = L.METHODS[method] if isinstance(method, str) else int(method)
    o.restart = int(restart or 0)
    o.rows_per_thread = int(rows_per_thread or 0)
    o.fold_colour0 = int(fold_colour0 or 0)
    return o


class DeviceMatrix:
    """kcfd_mat: canonical CSR + rhs + diagonal tracker on the device."""

    def __init__(self, ctx: L.Context, indptr, indices, handle=None):
        self.ctx = ctx
        self.indptr = _i32(indptr)
        self.indices = _i32(indices)
        self.n = len(self.indptr) - 1
        self.nnz = int(self.indptr[-1])
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().kcfd_mat_create(ctx.h, self.n, self.nnz, ptr(self.indptr), ptr(self.indices),
                                        ctypes.byref(h)))
            handle = h
        self.h = handle
        v, r, d = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().kcfd_mat_dev_ptrs(self.h, ctypes.byref(v), ctypes.byref(r), ctypes.byref(d)))
        self.vals_ptr, self.rhs_ptr, self.diag_ptr = v, r, d
        self.pc = None

    @classmethod
    def from_csr(cls, ctx: L.Context, A) -> "DeviceMatrix":
        A = sparse.csr_matrix(A)
        if not A.has_canonical_format:
            A = A.copy()
            A.sum_duplicates()
        m = cls(ctx, A.indptr, A.indices)
        m.set_values(A.data)
        return m

    # values ----------------------------------------------------------------
    def set_values(self, vals):
        v = _f64(vals)
        assert v.size == self.nnz
        check(lib().kcfd_mat_set_values(self.h, ptr(v)))
        self.pc = None

    def get_values(self) -> np.ndarray:
        out = np.empty(self.nnz)
        check(lib().kcfd_mat_get_values(self.h, ptr(out)))
        return out

    def _download(self, p, n) -> np.ndarray:
        out = np.empty(n)
        check(lib().kcfd_buf_download(self.ctx.h, ptr(out), p, 8 * n))
        return out

    def _upload(self, p, a):
        a = _f64(a)
        check(lib().kcfd_buf_upload(self.ctx.h, p, ptr(a), 8 * a.size))

    def get_rhs(self) -> np.ndarray:
        return self._download(self.rhs_ptr, self.n)

    def set_rhs(self, a):
        self._upload(self.rhs_ptr, a)

    def get_diag_tracker(self) -> np.ndarray:
        return self._download(self.diag_ptr, self.n)

    def set_diag_tracker(self, a):
        self._upload(self.diag_ptr, a)

    def to_scipy(self) -> sparse.csr_matrix:
        return sparse.csr_matrix((self.get_values(), self.indices.copy(), self.indptr.copy()),
                                 shape=(self.n, self.n))

    def bytes(self) -> int:
        b = ctypes.c_longlong()
        check(lib().kcfd_mat_bytes(self.h, ctypes.byref(b)))
        return b.value

    # solver ----------------------------------------------------------------
    def setup_precond(self, kind="jacobi", precision="fp64", order=None):
        """Set up the preconditioner from the current values.  ``order``
        ('natural' or 'color') selects the ILU(0) row ordering; None keeps
        the matrix's current one (natural by default)."""
        k = L.PRECONDS[kind] if isinstance(kind, str) else int(kind)
        p = {"fp64": L.FP64, "mixed": L.MIXED}[precision] if isinstance(precision, str) else int(precision)
        if order is not None:
            o = L.ORDERS[order] if isinstance(order, str) else int(order)
            check(lib().kcfd_precond_set_order(self.h, o))
        check(lib().kcfd_precond_setup(self.h, k, p))
        self.pc = (k, p)

    def ilu0_factors(self):
        """(lu on the pattern of P A P^T, levels of L, levels of U)."""
        lu = np.empty(self.nnz)
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().kcfd_ilu0_get(self.h, ptr(lu), ctypes.byref(a), ctypes.byref(b)))
        return lu, a.value, b.value

    def copy_order_from(self, other: "DeviceMatrix"):
        """Take ``other``'s ILU(0) ordering and level schedule (same structure)."""
        check(lib().kcfd_mat_copy_order(self.h, other.h))

    def ilu0_order(self):
        """(perm, ncolors) of the current ILU(0) factors: perm[k] is the row of
        A at position k of P A P^T (identity and 0 colours when natural)."""
        perm = np.empty(self.n, np.int32)
        nc = ctypes.c_int()
        check(lib().kcfd_ilu0_order(self.h, ptr(perm), ctypes.byref(nc)))
        return perm, nc.value

    def spmv(self, x_dev: DevBuf, y_dev: DevBuf):
        check(lib().kcfd_mat_spmv(self.h, x_dev.p, y_dev.p))

    def solve_dev(self, b_ptr, x_ptr, opts: L.SolveOpts, setup: bool = True,
                  fallback: bool = True) -> L.SolveResult:
        """Solve on device buffers.  The preconditioner is rebuilt from the
        current values (as create_preconditioner is called on every
        solve_linear_system, core/linear_solver.py:56) unless ``setup`` is
        False and it was set up for these options already.

        A mixed-precision solve that stops without converging (FP32 Krylov
        recurrences can stall on very ill-conditioned systems) continues in
        FP64 from the best iterate when ``fallback`` is set; the result then
        counts both phases and ``res.fallback`` is 1."""
        if setup or self.pc != (opts.precond, opts.precision):
            self.setup_precond(opts.precond, opts.precision)
        res = L.SolveResult()
        check(lib().kcfd_solve(self.h, b_ptr, x_ptr, ctypes.byref(opts), ctypes.byref(res)))
        res.fallback = 0
        if fallback and opts.precision == L.MIXED and res.info != 0:
            res = fp64_continuation(self, b_ptr, x_ptr, opts, res)
        return res

    def solve(self, b, x0=None, **kw):
        """Host convenience: returns (x, SolveResult)."""
        opts = default_opts(**kw)
        bb = DevBuf.from_array(self.ctx, b)
        xx = DevBuf.from_array(self.ctx, np.zeros(self.n) if x0 is None else x0)
        res = self.solve_dev(bb.p, xx.p, opts)
        return xx.download(), res

    def close(self):
        if getattr(self, "h", None):
            lib().kcfd_mat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mesh_structure(mesh):
    """Canonical CSR of diag U {(O,N),(N,O) : internal faces} and the static
    face->slot map (slot_on, slot_no) and diag_slot."""
    n = mesh.n_cells
    o = mesh.owner.astype(np.int64)
    nb = mesh.neighbour.astype(np.int64)
    internal = nb >= 0
    io, inb = o[internal], nb[internal]
    ar = np.arange(n, dtype=np.int64)
    keys = np.concatenate([ar * n + ar, io * n + inb, inb * n + io])
    uk = np.unique(keys)
    rows = uk // n
    cols = uk % n
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr)
    slot_on = np.full(mesh.n_faces, -1, np.int64)
    slot_no = np.full(mesh.n_faces, -1, np.int64)
    slot_on[internal] = np.searchsorted(uk, io * n + inb)
    slot_no[internal] = np.searchsorted(uk, inb * n + io)
    diag_slot = np.searchsorted(uk, ar * n + ar)
    return (indptr.astype(np.int32), cols.astype(np.int32), slot_on.astype(np.int32),
            slot_no.astype(np.int32), diag_slot.astype(np.int32))


class DeviceMesh:
    """kcfd_mesh: SoA geometry, cell->face CSR and the face->slot map."""

    def __init__(self, mesh, ctx: L.Context | None = None):
        self.ctx = ctx or L.default_context()
        self.mesh = mesh
        d = mesh.derived()
        self.indptr, self.indices, self.slot_on, self.slot_no, self.diag_slot = mesh_structure(mesh)
        names = list(mesh.boundary_patches.keys())
        self.patch_names = names
        offs = [0]
        for nm in names:
            offs.append(offs[-1] + len(mesh.boundary_patches[nm]))
        self.patch_offsets = offs
        self.n_bfaces = offs[-1]
        bidx = np.full(mesh.n_faces, -1, np.int64)
        for pi, nm in enumerate(names):
            f = np.asarray(mesh.boundary_patches[nm], np.int64)
            bidx[f] = offs[pi] + np.arange(len(f))
        self._keep = dict(
            owner=_i32(mesh.owner), neighbour=_i32(mesh.neighbour), area=_f64(mesh.face_area),
            normal=_f64(mesh.face_normal), fcenter=_f64(mesh.face_center),
            ccenter=_f64(mesh.cell_center), volume=_f64(mesh.cell_volume), d_pn=_f64(d["d_pn"]),
            d_pf=_f64(d["d_pf"]), gweight=_f64(d["gweight"]), bpatch=_i32(d["bpatch"]),
            bidx=_i32(bidx), cf_ptr=_i32(d["cf_ptr"]), cf_face=_i32(d["cf_face"]),
            cf_side=np.ascontiguousarray(d["cf_side"], np.int8), indptr=self.indptr,
            indices=self.indices, slot_on=self.slot_on, slot_no=self.slot_no,
            diag_slot=self.diag_slot)
        desc = L.MeshDesc()
        desc.n_cells = mesh.n_cells
        desc.n_faces = mesh.n_faces
        desc.ndim = mesh.ndim
        desc.n_patches = len(names)
        desc.nnz = int(self.indptr[-1])
        for k, v in self._keep.items():
            setattr(desc, k, v.ctypes.data)
        h = ctypes.c_void_p()
        check(lib().kcfd_mesh_create(self.ctx.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        self.n = mesh.n_cells
        self.nf = mesh.n_faces
        self.ndim = mesh.ndim
        self.nnz = int(self.indptr[-1])

    def new_system(self) -> DeviceMatrix:
        h = ctypes.c_void_p()
        check(lib().kcfd_mesh_new_system(self.h, ctypes.byref(h)))
        return DeviceMatrix(self.ctx, self.indptr, self.indices, handle=h)

    def boundary_array(self, values: dict, width: int = 1, default=None) -> np.ndarray:
        """Concatenate per-patch boundary arrays in patch order."""
        parts = []
        for nm in self.patch_names:
            cnt = len(self.mesh.boundary_patches[nm])
            if nm in values:
                a = np.asarray(values[nm], np.float64).reshape(cnt, width) if width > 1 else \
                    np.asarray(values[nm], np.float64).reshape(cnt)
            else:
                a = np.zeros((cnt, width) if width > 1 else cnt) if default is None else default(nm)
            parts.append(a)
        if not parts:
            return np.zeros((0, width) if width > 1 else 0)
        return np.concatenate(parts)

    def last_ms(self) -> float:
        v = ctypes.c_double()
        check(lib().kcfd_last_kernel_ms(self.ctx.h, ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().kcfd_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """kcfd_batch: several device-resident systems solved by the same launches."""

    def __init__(self, ctx: L.Context, mats, bs, xs, opts: L.SolveOpts):
        self.ctx = ctx
        self.mats = list(mats)
        self.bs = list(bs)
        self.xs = list(xs)
        n = len(self.mats)
        MA = ctypes.c_void_p * n
        self._m = MA(*[m.h.value for m in self.mats])
        self._b = MA(*[b.p.value if isinstance(b, DevBuf) else int(b) for b in self.bs])
        self._x = MA(*[x.p.value if isinstance(x, DevBuf) else int(x) for x in self.xs])
        self.opts = opts
        h = ctypes.c_void_p()
        check(lib().kcfd_batch_create(ctx.h, n, ctypes.cast(self._m, ctypes.c_void_p),
                                      ctypes.cast(self._b, ctypes.c_void_p),
                                      ctypes.cast(self._x, ctypes.c_void_p), ctypes.byref(opts),
                                      ctypes.byref(h)))
        self.h = h
        self.nsys = n

    def run(self):
        res = (L.SolveResult * self.nsys)()
        check(lib().kcfd_batch_run(self.h, ctypes.cast(res, ctypes.c_void_p)))
        return list(res)

    def bytes(self) -> int:
        b = ctypes.c_longlong()
        check(lib().kcfd_batch_bytes(self.h, ctypes.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "h", None):
            lib().kcfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
