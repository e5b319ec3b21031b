"""ctypes binding of libkcfd.so (declarations in include/kcfd.h).

argtypes/restypes are declared once here.  The library is the only compute
path: if it is missing or no GPU is present, calls raise ``KcfdError`` (there
is no CPU fallback).
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkcfd.so")

KCFD_OK = 0
KCFD_EARG, KCFD_EHIP, KCFD_EOOM, KCFD_ESTATE, KCFD_EPIVOT, KCFD_EBUSY = -1, -2, -3, -4, -5, -6
PC_NONE, PC_JACOBI, PC_ILU0 = 0, 1, 2
FP64, MIXED = 0, 1
METHOD_BICGSTAB, METHOD_CG, METHOD_GMRES = 0, 1, 2
METHODS = {"bicgstab": METHOD_BICGSTAB, "cg": METHOD_CG, "gmres": METHOD_GMRES}
BC_ZERO_GRADIENT, BC_DIRICHLET, BC_NEUMANN, BC_NONE = 0, 1, 2, 3
LIMITERS = {"minmod": 0, "superbee": 1, "van_albada": 2, "van_leer": 3}
PRECONDS = {"none": PC_NONE, "jacobi": PC_JACOBI, "ilu0": PC_ILU0}
ORDER_NATURAL, ORDER_COLOR, ORDER_SEGMENT = 0, 1, 2
ORDERS = {"natural": ORDER_NATURAL, "color": ORDER_COLOR, "segment": ORDER_SEGMENT}

_ERRNAMES = {KCFD_EARG: "bad argument", KCFD_EHIP: "HIP error", KCFD_EOOM: "out of device memory",
             KCFD_ESTATE: "bad state", KCFD_EPIVOT: "zero pivot", KCFD_EBUSY: "admission refused"}


class KcfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libkcfd {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class SolveOpts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("maxiter", ctypes.c_int),
                ("precond", ctypes.c_int), ("precision", ctypes.c_int),
                ("inner_rtol", ctypes.c_double), ("max_refine", ctypes.c_int),
                ("check_every", ctypes.c_int), ("grid_per_cu", ctypes.c_int),
                ("trsv_per_cu", ctypes.c_int), ("tile_walk", ctypes.c_int),
                ("method", ctypes.c_int), ("restart", ctypes.c_int), ("rows_per_thread", ctypes.c_int),
                ("fold_colour0", ctypes.c_int)]


class SolveResult(ctypes.Structure):
    _fields_ = [("info", ctypes.c_int), ("iters", ctypes.c_int), ("half_step", ctypes.c_int),
                ("refinements", ctypes.c_int), ("rel_resid", ctypes.c_double),
                ("bytes", ctypes.c_double), ("ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["fallback"] = getattr(self, "fallback", 0)   # mixed solve finished in FP64
        return d


_P = ctypes.c_void_p


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int), ("n_faces", ctypes.c_int), ("ndim", ctypes.c_int),
                ("n_patches", ctypes.c_int), ("nnz", ctypes.c_int),
                ("owner", _P), ("neighbour", _P), ("area", _P), ("normal", _P), ("fcenter", _P),
                ("ccenter", _P), ("volume", _P), ("d_pn", _P), ("d_pf", _P), ("gweight", _P),
                ("bpatch", _P), ("bidx", _P), ("cf_ptr", _P), ("cf_face", _P), ("cf_side", _P),
                ("indptr", _P), ("indices", _P), ("slot_on", _P), ("slot_no", _P),
                ("diag_slot", _P)]


# name -> (argtypes, restype); every symbol declared in include/kcfd.h
_I, _D, _SZ = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "kcfd_init": ([_I, _PP], _I),
    "kcfd_destroy": ([_P], _I),
    "kcfd_last_error": ([], ctypes.c_char_p),
    "kcfd_device_info": ([_P, ctypes.c_char_p, _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong)], _I),
    "kcfd_sync": ([_P], _I),
    "kcfd_stream_copy": ([_P, _SZ, _I, ctypes.POINTER(_D)], _I),
    "kcfd_stream_read": ([_P, _SZ, _I, ctypes.POINTER(_D)], _I),
    "kcfd_buf_alloc": ([_P, _SZ, _PP], _I),
    "kcfd_buf_free": ([_P, _P], _I),
    "kcfd_buf_upload": ([_P, _P, _P, _SZ], _I),
    "kcfd_buf_download": ([_P, _P, _P, _SZ], _I),
    "kcfd_buf_fill": ([_P, _P, _D, _SZ], _I),
    "kcfd_mat_create": ([_P, _I, _I, _P, _P, _PP], _I),
    "kcfd_mat_destroy": ([_P], _I),
    "kcfd_mat_set_values": ([_P, _P], _I),
    "kcfd_mat_get_values": ([_P, _P], _I),
    "kcfd_mat_dev_ptrs": ([_P, _PP, _PP, _PP], _I),
    "kcfd_mat_spmv": ([_P, _P, _P], _I),
    "kcfd_mat_bytes": ([_P, ctypes.POINTER(ctypes.c_longlong)], _I),
    "kcfd_precond_setup": ([_P, _I, _I], _I),
    "kcfd_ilu0_get": ([_P, _P, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "kcfd_precond_set_order": ([_P, _I], _I),
    "kcfd_mat_copy_order": ([_P, _P], _I),
    "kcfd_batch_create_dynamic": ([_P, _I, ctypes.POINTER(SolveOpts), _PP], _I),
    "kcfd_batch_attach": ([_P, _P, _P, _P, ctypes.POINTER(_I)], _I),
    "kcfd_batch_step": ([_P, _P, ctypes.POINTER(_I)], _I),
    "kcfd_batch_detach": ([_P, _I, ctypes.POINTER(SolveResult)], _I),
    "kcfd_batch_members": ([_P, ctypes.POINTER(_I)], _I),
    "kcfd_batch_trim": ([_P], _I),
    "kcfd_ilu0_order": ([_P, _P, ctypes.POINTER(_I)], _I),
    "kcfd_precond_apply": ([_P, _P, _P], _I),
    "kcfd_solve": ([_P, _P, _P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveResult)], _I),
    "kcfd_schiller_naumann": ([_P, _P, _P, ctypes.c_longlong], _I),
    "kcfd_drag_implicit": ([_P, _I, _I, _P, _D, _P, _P, _D, _D, _P], _I),
    "kcfd_solve_noalloc": ([_P, _P, _P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveResult)], _I),
    "kcfd_batch_create": ([_P, _I, _P, _P, _P, ctypes.POINTER(SolveOpts), _PP], _I),
    "kcfd_batch_run": ([_P, _P], _I),
    "kcfd_batch_destroy": ([_P], _I),
    "kcfd_batch_bytes": ([_P, ctypes.POINTER(ctypes.c_longlong)], _I),
    "kcfd_mat_solver": ([_P, _PP], _I),
    "kcfd_batch_profile": ([_P, _I, _I, ctypes.POINTER(_D)], _I),
    "kcfd_batch_kernel_bytes": ([_P, _I, ctypes.POINTER(_D)], _I),
    "kcfd_mesh_build": ([_P, _I, _I, _P, _I, _I, _P, _I, _I, _P, _P, _I, _P, _P, _I, _PP], _I),
    "kcfd_mesh_faces_info": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                              ctypes.POINTER(_D)], _I),
    "kcfd_mesh_faces_get": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "kcfd_mesh_faces_free": ([_P], _I),
    "kcfd_mesh_generate": ([_P, _I, _I, _I, _I, _D, _D, _D, _D, _D, _P, _I, _PP], _I),
    "kcfd_mesh_faces_nodes": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), _P], _I),
    "kcfd_mesh_create": ([_P, ctypes.POINTER(MeshDesc), _PP], _I),
    "kcfd_mesh_destroy": ([_P], _I),
    "kcfd_mesh_new_system": ([_P, _PP], _I),
    "kcfd_asm_reset": ([_P], _I),
    "kcfd_asm_diffusion": ([_P, _P, _P], _I),
    "kcfd_asm_upwind": ([_P, _P, _P], _I),
    "kcfd_asm_boundary": ([_P, _P, _P, _P, _P, _P, _P], _I),
    "kcfd_asm_temporal": ([_P, _P, _D, _D, _P], _I),
    "kcfd_asm_source": ([_P, _P, _P], _I),
    "kcfd_asm_linearized": ([_P, _P, _P, _P], _I),
    "kcfd_asm_cell_terms": ([_P, _P, _P, _P], _I),
    "kcfd_asm_under_relax": ([_P, _P, _D], _I),
    "kcfd_asm_pressure": ([_P, _P, _P, _D, _P], _I),
    "kcfd_grad_gg": ([_P, _P, _P, _P], _I),
    "kcfd_grad_lsq": ([_P, _P, _P, _P], _I),
    "kcfd_simple_dp": ([_P, _P, _P, _P, _P], _I),
    "kcfd_simple_rhs": ([_P, _P, _P], _I),
    "kcfd_simple_fix_reference": ([_P], _I),
    "kcfd_simple_correct": ([_P, _P, _P, _P, ctypes.c_double, _P, _P], _I),
    "kcfd_twofluid_pressure": ([_P, _P, _P, _P, _P, _P, _D, _D, _P, _P], _I),
    "kcfd_twofluid_correct": ([_P, _P, _P, _P, _P, _P, _D, _P, _P, _P], _I),
    "kcfd_muscl": ([_P, _P, _P, _P, _I, _P], _I),
    "kcfd_face_values": ([_P, _I, _P, _P, _P, _P, _P], _I),
    "kcfd_limiter": ([_P, _I, _P, _P, ctypes.c_longlong], _I),
    "kcfd_mass_flux": ([_P, _P, _P, _D, _P], _I),
    "kcfd_last_kernel_ms": ([_P, ctypes.POINTER(_D)], _I),
}

_lib = None
_lock = threading.RLock()


def lib():
    """Load libkcfd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise KcfdError(KCFD_ESTATE, f"{LIB_PATH} is not built; run "
                                    "`python k-cfd_amd/build.py` (hipcc, gfx950)")
                L = ctypes.CDLL(LIB_PATH)
                for name, (args, res) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = res
                _lib = L
    return _lib


def check(rc: int):
    if rc != KCFD_OK:
        raise KcfdError(rc, lib().kcfd_last_error().decode(errors="replace"))


def ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One device + one stream (kcfd_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().kcfd_init(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cus = ctypes.c_int()
        tot = ctypes.c_longlong()
        fr = ctypes.c_longlong()
        check(lib().kcfd_device_info(self.h, name, 256, ctypes.byref(cus), ctypes.byref(tot),
                                     ctypes.byref(fr)))
        return {"name": name.value.decode(), "cus": cus.value, "total_bytes": tot.value,
                "free_bytes": fr.value}

    def sync(self):
        check(lib().kcfd_sync(self.h))

    def stream_copy_gbs(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Measured device copy bandwidth (read + write bytes / time)."""
        ms = ctypes.c_double()
        check(lib().kcfd_stream_copy(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return 2.0 * nbytes / (ms.value * 1e-3) / 1e9

    def stream_read_gbs(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Measured device read bandwidth (bytes read / time)."""
        ms = ctypes.c_double()
        check(lib().kcfd_stream_read(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return nbytes / (ms.value * 1e-3) / 1e9

    def close(self):
        if getattr(self, "h", None):
            lib().kcfd_destroy(self.h)
            self.h = None


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        with _lock:
            if _default_ctx is None:
                _default_ctx = Context(int(os.environ.get("KCFD_DEVICE", "0")))
    return _default_ctx


class DevBuf:
    """A device array owned by the library's allocator."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().kcfd_buf_alloc(ctx.h, max(self.nbytes, 8), ctypes.byref(p)))
        self.p = p

    @classmethod
    def from_array(cls, ctx: Context, a, dtype=np.float64) -> "DevBuf":
        a = np.ascontiguousarray(a, dtype=dtype)
        b = cls(ctx, a.size, dtype)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch: buffer {self.n}, array {a.size}")
        check(lib().kcfd_buf_upload(self.ctx.h, self.p, ptr(a), self.nbytes))

    def download(self, out=None) -> np.ndarray:
        out = np.empty(self.n, self.dtype) if out is None else out
        check(lib().kcfd_buf_download(self.ctx.h, ptr(out), self.p, self.nbytes))
        return out

    def fill(self, value: float):
        check(lib().kcfd_buf_fill(self.ctx.h, self.p, float(value), self.n))

    def free(self):
        if getattr(self, "p", None) and self.p.value:
            try:
                lib().kcfd_buf_free(self.ctx.h, self.p)
            finally:
                self.p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
