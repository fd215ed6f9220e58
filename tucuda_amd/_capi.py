"""ctypes binding of libtucuda_amd.so (the C ABI declared in include/tucu.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtucuda_amd.so")

TUCU_OK = 0
TUCU_ERR_INVALID = 1
TUCU_ERR_LOBE_TOO_LARGE = 2
TUCU_ERR_UNSUPPORTED_LOBE = 3
TUCU_ERR_BAD_ALLOC = 4
TUCU_ERR_DEVICE = 5

TUCU_U8, TUCU_F32, TUCU_F16, TUCU_BF16 = 0, 1, 2, 3
TUCU_RULE_REFERENCE, TUCU_RULE_CEIL4SIGMA = 0, 1
TUCU_ALL6 = 0x3F
TUCU_ICIA_AFFINE, TUCU_ICIA_PROJECTIVE = 6, 8

# cu::subsample operators (functional.h:58-256), TUCU_OP_* in include/tucu.h.
SUBSAMPLE_OPS = ("identity1x1", "diffH1x3", "diffV3x1", "diffHH1x3", "diffVV3x1", "diffHV3x3",
                 "sobelH3x3", "sobelV3x3", "sobelAbs3x3", "laplacian3x3", "det3x3", "gaussian5x5")

# Every symbol include/tucu.h declares.
EXPORTS = (
    "tucu_last_error", "tucu_version", "tucu_malloc", "tucu_free", "tucu_row_stride",
    "tucu_memcpy2d", "tucu_synchronize", "tucu_fir2_check", "tucu_fir2_convolve",
    "tucu_fir2_convolve_taps",
    "tucu_lobe_size", "tucu_subsample", "tucu_gauss2_pyramid", "tucu_gauss2_lobes", "tucu_gauss2_convolve",
    "tucu_icia_moment",
    "tucu_batch_create", "tucu_batch_destroy", "tucu_batch_launch", "tucu_batch_set_tile_rows",
    "tucu_batch_autotune", "tucu_batch_info", "tucu_batch_run_static", "tucu_tune_cache_save",
    "tucu_tune_cache_load",
    "tucu_server_admit_bytes", "tucu_server_create", "tucu_server_destroy", "tucu_server_now", "tucu_server_submit",
    "tucu_server_submit_many", "tucu_server_poll", "tucu_server_release", "tucu_server_stats",
)


class TucuImage(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("stride", ctypes.c_int64),
        ("nrow", ctypes.c_int32),
        ("ncol", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class TucuGaussJob(ctypes.Structure):
    _fields_ = [
        ("in_", TucuImage),
        ("out", TucuImage * 6),
        ("lobe0", ctypes.POINTER(ctypes.c_float)),
        ("lobe1", ctypes.POINTER(ctypes.c_float)),
        ("lobe2", ctypes.POINTER(ctypes.c_float)),
        ("n0", ctypes.c_int32),
        ("n1", ctypes.c_int32),
        ("n2", ctypes.c_int32),
        ("mask", ctypes.c_uint32),
        ("shift", ctypes.c_int32),
    ]


class TucuServerConfig(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_size_t),
        ("in_dtype", ctypes.c_int32),
        ("out_dtype", ctypes.c_int32),
        ("rule", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("max_batch_jobs", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("max_tiles", ctypes.c_int32),
        ("max_batch_mpix", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("max_pool_bytes", ctypes.c_size_t),
        ("reserve_bytes", ctypes.c_size_t),
        ("grow_bytes", ctypes.c_size_t),
    ]


class TucuRequest(ctypes.Structure):
    _fields_ = [
        ("in_", TucuImage),
        ("sigma", ctypes.c_float),
        ("mask", ctypes.c_uint32),
        ("shift", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("arrival", ctypes.c_double),
    ]


class TucuCompletion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("arrival", ctypes.c_double),
        ("admitted", ctypes.c_double),
        ("launched", ctypes.c_double),
        ("completed", ctypes.c_double),
        ("out", TucuImage * 6),
    ]


class TucuServerCounters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "batches", "requests_launched", "requests_done", "max_batch",
        "pool_bytes", "pool_used", "pool_peak")]


class TucuError(RuntimeError):
    """A failure reported by the C ABI; ``code`` is the tucu_status."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class LobeTooLarge(TucuError):
    """FIRFilter.h:250-251 (runtime_error at initialize)."""


class UnsupportedLobe(TucuError):
    """FIRFilter.h:343-344, 374-375 (runtime_error at convolve)."""


class BadAlloc(TucuError, MemoryError):
    """allocator.h:83-84 (std::bad_alloc)."""


class DeviceError(TucuError):
    """A HIP runtime failure."""


_ERRORS = {
    TUCU_ERR_LOBE_TOO_LARGE: LobeTooLarge,
    TUCU_ERR_UNSUPPORTED_LOBE: UnsupportedLobe,
    TUCU_ERR_BAD_ALLOC: BadAlloc,
    TUCU_ERR_DEVICE: DeviceError,
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or make -C tucuda_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    imgp = ctypes.POINTER(TucuImage)
    L.tucu_last_error.restype = ctypes.c_char_p
    L.tucu_last_error.argtypes = []
    L.tucu_version.restype = ctypes.c_int
    L.tucu_malloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.tucu_free.argtypes = [ctypes.c_void_p]
    L.tucu_row_stride.restype = ctypes.c_size_t
    L.tucu_row_stride.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    L.tucu_memcpy2d.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    L.tucu_synchronize.argtypes = []
    L.tucu_fir2_check.argtypes = [ctypes.c_int, ctypes.c_int]
    L.tucu_fir2_convolve.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, imgp, imgp,
                                     ctypes.c_int, ctypes.c_void_p]
    L.tucu_fir2_convolve_taps.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, imgp, imgp,
                                          ctypes.c_int, ctypes.c_void_p]
    L.tucu_lobe_size.argtypes = [fp, ctypes.c_int]
    L.tucu_subsample.argtypes = [imgp, imgp, ctypes.c_int, ctypes.c_void_p]
    L.tucu_gauss2_pyramid.argtypes = [fp, ctypes.c_int, imgp, imgp, ctypes.c_int, ctypes.c_void_p]
    L.tucu_gauss2_lobes.argtypes = [ctypes.c_float, ctypes.c_int, fp, ip, fp, ip, fp, ip]
    L.tucu_gauss2_convolve.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, fp, ctypes.c_int,
                                       imgp, imgp, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p]
    vp = ctypes.c_void_p
    L.tucu_icia_moment.argtypes = [fp, ctypes.c_int, fp, ctypes.c_int, imgp, ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
    L.tucu_server_admit_bytes.argtypes = [ctypes.c_size_t] * 6
    L.tucu_server_admit_bytes.restype = ctypes.c_size_t
    L.tucu_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(TucuGaussJob), ctypes.c_int]
    L.tucu_batch_destroy.argtypes = [vp]
    L.tucu_batch_launch.argtypes = [vp, vp]
    L.tucu_batch_run_static.argtypes = [ctypes.POINTER(TucuGaussJob), ctypes.c_int, vp]
    L.tucu_batch_set_tile_rows.argtypes = [vp, ctypes.c_int]
    L.tucu_batch_autotune.argtypes = [vp, vp, ctypes.c_int, ip, fp]
    L.tucu_batch_info.argtypes = [vp, ip, ip, ip, ip]
    L.tucu_tune_cache_save.argtypes = [ctypes.c_char_p]
    L.tucu_tune_cache_load.argtypes = [ctypes.c_char_p]
    L.tucu_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(TucuServerConfig)]
    L.tucu_server_destroy.argtypes = [vp]
    L.tucu_server_now.argtypes = [vp]
    L.tucu_server_now.restype = ctypes.c_double
    L.tucu_server_submit.argtypes = [vp, ctypes.POINTER(TucuRequest), ctypes.POINTER(ctypes.c_uint64)]
    L.tucu_server_submit_many.argtypes = [vp, ctypes.POINTER(TucuRequest), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_uint64)]
    L.tucu_server_poll.argtypes = [vp, ctypes.POINTER(TucuCompletion), ctypes.c_int, ip]
    L.tucu_server_release.argtypes = [vp, ctypes.c_uint64]
    L.tucu_server_stats.argtypes = [vp, ctypes.POINTER(TucuServerCounters)]
    for name in EXPORTS:
        if name not in ("tucu_last_error", "tucu_row_stride", "tucu_version", "tucu_server_now",
                        "tucu_server_admit_bytes"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    """Raise the exception matching a tucu_status."""
    if rc == TUCU_OK:
        return
    msg = load().tucu_last_error().decode(errors="replace")
    raise _ERRORS.get(rc, TucuError)(rc, msg)
