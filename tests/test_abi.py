"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/kcfd.h declares, the ctypes declarations match, and the entry points
fail loudly without a GPU (no CPU fallback)."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "kcfd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kcfd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_all_bound_symbols():
    from kcfd_amd import _lib
    decl = set(_declared())
    assert decl == set(_lib.SIGNATURES), (decl ^ set(_lib.SIGNATURES))


def test_library_exports_every_declared_symbol():
    from kcfd_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "run python k-cfd_amd/build.py first"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(kcfd_[a-z0-9_]+)", out))
    missing = set(_declared()) - exported
    assert not missing, missing
    L = _lib.lib()
    for name in _declared():
        assert getattr(L, name) is not None


def test_header_compiles_as_c():
    src = '#include "kcfd.h"\nint main(void){kcfd_solve_opts o; (void)o; return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-", "-o", "/dev/null"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_header():
    """ctypes structure sizes equal the C sizes."""
    from kcfd_amd import _lib
    src = ('#include "kcfd.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu\\n",'
           'sizeof(kcfd_solve_opts),sizeof(kcfd_solve_result),sizeof(kcfd_mesh_desc));return 0;}\n')
    exe = "/tmp/kcfd_sizes"
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-x", "c", "-", "-o", exe],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    import ctypes
    assert sizes == [ctypes.sizeof(_lib.SolveOpts), ctypes.sizeof(_lib.SolveResult),
                     ctypes.sizeof(_lib.MeshDesc)]


def test_no_gpu_fails_loudly():
    """Without a GPU the library raises instead of computing on the CPU."""
    import ctypes
    from kcfd_amd import _lib
    h = ctypes.c_void_p()
    rc = _lib.lib().kcfd_init(0, ctypes.byref(h))
    if rc == 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.KcfdError):
        _lib.check(rc)
    assert _lib.lib().kcfd_last_error()


def test_argument_validation_without_gpu():
    """Null handles are rejected with KCFD_EARG before any device call."""
    import ctypes
    from kcfd_amd import _lib
    L = _lib.lib()
    assert L.kcfd_mat_spmv(None, None, None) == _lib.KCFD_EARG
    assert L.kcfd_solve(None, None, None, None, None) == _lib.KCFD_EARG
    assert L.kcfd_batch_run(None, None) == _lib.KCFD_EARG
    assert b"invalid argument" in L.kcfd_last_error()


def test_default_context_raises_without_gpu():
    """default_context() raises (no deadlock, no fallback) when no device is visible."""
    from kcfd_amd import _lib
    import ctypes
    h = ctypes.c_void_p()
    if _lib.lib().kcfd_init(0, ctypes.byref(h)) == 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.KcfdError):
        _lib.default_context()
