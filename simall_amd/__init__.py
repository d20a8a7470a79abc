"""MI355X (gfx950) backend for the SimAll pressure-correction solve path.

The product is the C-ABI library ``libsimall_amd.so`` (include/simall_amd.h):
hand-written HIP kernels for SpMV, fused Krylov iterations, preconditioners
and finite-volume assembly.  ``simall_amd.capi`` is a thin ctypes binding used
by the tests and the benchmark; it raises when the library is missing.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
