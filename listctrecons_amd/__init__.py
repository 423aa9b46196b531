"""MI355X-native helical cone-beam FDK (LISTCTRecons hot path).

The product is liblct.so (HIP kernels for gfx950 behind the C ABI in
include/lct.h), the C++ drop-in classes over it (listctrecons_amd/cpp) and
this thin Python host layer used by the tests and the benchmark.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib", "scan", "recon"]
