"""hypre_amd -- MI355X-native BoomerAMG-preconditioned PCG behind hypre's C API.

The product is the C-ABI shared library ``hypre_amd/libhypre_mi355x.so``
(declared in ``include/*.h``).  This package only loads it through ctypes and
offers thin Python wrappers for tests and benchmarks; every compute call goes
to the HIP kernels in the library.  There is no CPU fallback: loading fails
loudly when the library is missing.
"""
from .lib import LIB_PATH, lib  # noqa: F401
from . import api  # noqa: F401
