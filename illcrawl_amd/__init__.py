"""illcrawl_amd: MI355X-native projective SPH smoothing (illcrawl's hot path).

The device work lives in libilcpsg.so (HIP, gfx950) behind the C ABI of
include/illcrawl_psg.h.  This package holds the Python mirror of the
reference's backend interface (``backend``), the synthetic workloads
(``synth``) and the build script (``build``).
"""
from ._native import IllcrawlError, NativeLibraryMissing, LIB_PATH  # noqa: F401

__all__ = ["IllcrawlError", "NativeLibraryMissing", "LIB_PATH"]
