"""tucuda_amd -- MI355X-native separable FIR / Gaussian-derivative filtering.

A from-scratch gfx950 implementation of TUCuda's FIRFilter2 /
FIRGaussianConvolver2 path.  The compute lives in ``libtucuda_amd.so``
(hand-written HIP kernels behind a C ABI, ``include/tucu.h``); this package is
the Python mirror of the reference's C++ interface over that ABI, used by the
tests and the benchmark.  ``include/TU/cu/*.h`` is the C++ drop-in.
"""
from ._capi import (  # noqa: F401
    BadAlloc, DeviceError, LobeTooLarge, TucuError, UnsupportedLobe,
    TUCU_RULE_CEIL4SIGMA, TUCU_RULE_REFERENCE, LIB_PATH,
)


def __getattr__(name):
    # torch is imported lazily so that the ABI can be inspected without it.
    if name in ("FIRFilter2", "FIRGaussianConvolver2", "SeparableFilter2", "GaussBatch", "gauss_lobes",
                "lobe_size", "tune_cache_save", "tune_cache_load",
                "METHOD_ORDER", "METHOD_BITS", "METHOD_LOBES", "image_desc", "batch_tile_rows", "array2", "subsample",
                "SUBSAMPLE_OPS", "icia_moment", "ICIA_MODELS"):
        from . import filters
        return getattr(filters, name)
    raise AttributeError(name)
