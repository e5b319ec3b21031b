"""kcfd_amd -- MI355X-native FVM assembly + BiCGSTAB hot path of K-CFD.

The directory is ``k-cfd_amd``; it is imported as the package ``kcfd_amd``
(``__graft_entry__.load_package()`` registers it).  Layout:

  csrc/        HIP kernels for gfx950 and the C-ABI (include/kcfd.h)
  _lib.py      ctypes binding of libkcfd.so
  mesh.py      SoA mesh + vectorised builders (reference mesh rules)
  device.py    device-resident matrices, meshes, solves, batches
  core/        drop-in mirror of the reference's ``core`` API
  workloads.py seeded synthetic configurations
  batch.py     batched solver driver + autotuner
  serve.py     continuous-batching request loop
"""

from . import _lib  # noqa: F401
from ._lib import KcfdError, default_context  # noqa: F401

__all__ = ["KcfdError", "default_context"]
