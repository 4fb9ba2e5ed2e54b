"""MI355X-native 2D elastic P-SV FWI hot path.

The product is the native library cuda_fwi_amd/lib/libpsvfwi.so (hand-written
HIP kernels for gfx950 behind the C ABI in include/psvfwi.h, plus the C++
drop-in host layer).  ``psvfwi`` is its ctypes binding; ``autograd`` wraps it
as a torch.autograd.Function for differentiation tests.
"""
from . import psvfwi  # noqa: F401

__all__ = ["psvfwi"]
