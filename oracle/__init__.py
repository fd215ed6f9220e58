"""CPU oracle for TUCuda's separable FIR / Gaussian-derivative path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline) -- never as the thing measured or shipped.
The product library ``tucuda_amd`` does not import it.

This is a ctypes front end over ``liboracle.so`` (built from ``oracle.cpp`` by
``oracle/Makefile``).  ``oracle.cpp``'s header lists the reference lines each
function restates.  Image convention: numpy ``uint8`` = u_char, ``float32`` =
float, ``float16`` = binary16 storage, ``uint16`` = bfloat16 bit patterns.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

DT_U8, DT_F32, DT_F16, DT_BF16 = 0, 1, 2, 3
LOBE_SIZE_MAX = 17                      # FIRFilter.h:26
SUPPORTED_LOBES = (2, 3, 4, 5, 8, 9, 16, 17)   # FIRFilter.h:317-376

# (horizontal lobe, vertical lobe) of each FIRGaussianConvolver2 method,
# FIRGaussianConvolver.h:181, 193, 205, 217, 229, 241.
METHODS = {
    "smooth": (0, 0),
    "diffH": (1, 0),
    "diffV": (0, 1),
    "diffHH": (2, 0),
    "diffHV": (1, 1),
    "diffVV": (0, 2),
}
METHOD_ORDER = ("smooth", "diffH", "diffV", "diffHH", "diffHV", "diffVV")

# cu::subsample's stencil operators (functional.h:58-256), in code order.
SUBSAMPLE_OPS = ("identity1x1", "diffH1x3", "diffV3x1", "diffHH1x3", "diffVV3x1", "diffHV3x3",
                 "sobelH3x3", "sobelV3x3", "sobelAbs3x3", "laplacian3x3", "det3x3", "gaussian5x5")


class OracleError(RuntimeError):
    pass


class LobeTooLarge(OracleError):
    pass


class UnsupportedLobe(OracleError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise OracleError(f"{_LIB_PATH} is not built (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_lobe_size_f32.argtypes = [f32p, ctypes.c_int]
        L.orc_lobe_size_f64.argtypes = [f64p, ctypes.c_int]
        L.orc_gauss_lobes_f32.argtypes = [ctypes.c_float, f32p, ip, f32p, ip, f32p, ip]
        L.orc_gauss_lobes_f32.restype = None
        L.orc_gauss_lobes_f64.argtypes = [ctypes.c_double, f64p, ip, f64p, ip, f64p, ip]
        L.orc_gauss_lobes_f64.restype = None
        L.orc_gauss_lobes_ceil4_f32.argtypes = [ctypes.c_float, f32p, ip, f32p, ip, f32p, ip]
        L.orc_gauss_lobes_ceil4_f32.restype = None
        L.orc_gauss_lobes_ceil4_f64.argtypes = [ctypes.c_double, f64p, ip, f64p, ip, f64p, ip]
        L.orc_gauss_lobes_ceil4_f64.restype = None
        common = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.orc_fir2_f32.argtypes = [f32p, ctypes.c_int, f32p, ctypes.c_int] + common
        L.orc_fir2_f64.argtypes = [f64p, ctypes.c_int, f64p, ctypes.c_int] + common
        L.orc_fir2_taps_f32.argtypes = [f32p, ctypes.c_int, f32p, ctypes.c_int] + common
        L.orc_fir2_taps_f64.argtypes = [f64p, ctypes.c_int, f64p, ctypes.c_int] + common
        L.orc_subsample_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
        L.orc_icia_moment.argtypes = [f32p, f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      f64p, f64p]
        L.orc_half_to_float.argtypes = [ctypes.c_uint16]
        L.orc_half_to_float.restype = ctypes.c_float
        L.orc_float_to_half.argtypes = [ctypes.c_float]
        L.orc_float_to_half.restype = ctypes.c_uint16
        L.orc_bf16_to_float.argtypes = [ctypes.c_uint16]
        L.orc_bf16_to_float.restype = ctypes.c_float
        L.orc_float_to_bf16.argtypes = [ctypes.c_float]
        L.orc_float_to_bf16.restype = ctypes.c_uint16
        L.orc_float_to_u8.argtypes = [ctypes.c_float]
        L.orc_float_to_u8.restype = ctypes.c_uint8
        _lib = L
    return _lib


def _dt(a: np.ndarray) -> int:
    if a.dtype == np.uint8:
        return DT_U8
    if a.dtype == np.float32:
        return DT_F32
    if a.dtype == np.float16:
        return DT_F16
    if a.dtype == np.uint16:
        return DT_BF16
    raise TypeError(f"unsupported image dtype {a.dtype}")


def lobe_size(lobe17, even: bool, precision: str = "f32") -> int:
    """detail::lobeSize (FIRGaussianConvolver.cu:52-98)."""
    if precision == "f32":
        a = np.ascontiguousarray(lobe17, dtype=np.float32)
        assert a.size == LOBE_SIZE_MAX
        return lib().orc_lobe_size_f32(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(even))
    a = np.ascontiguousarray(lobe17, dtype=np.float64)
    assert a.size == LOBE_SIZE_MAX
    return lib().orc_lobe_size_f64(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(even))


def gauss_lobes(sigma: float, precision: str = "f32", rule: str = "reference"):
    """FIRGaussianConvolver2::initialize(sigma, lobe0, lobe1, lobe2)
    (FIRGaussianConvolver.h:123-170).  Returns three numpy arrays.

    ``rule`` "reference" sizes the lobes with detail::lobeSize
    (FIRGaussianConvolver.cu:52-98); "ceil4" fixes the radius at
    r = ceil(4 sigma), clamped to [1, 16] (lobes of r + 1, r, r + 1 entries),
    with the same normalisation formulas."""
    if rule not in ("reference", "ceil4"):
        raise ValueError(f"unknown lobe rule {rule!r}")
    c4 = rule == "ceil4"
    if precision == "f32":
        dt, ct = np.float32, ctypes.c_float
        fn = lib().orc_gauss_lobes_ceil4_f32 if c4 else lib().orc_gauss_lobes_f32
    else:
        dt, ct = np.float64, ctypes.c_double
        fn = lib().orc_gauss_lobes_ceil4_f64 if c4 else lib().orc_gauss_lobes_f64
    bufs = [np.zeros(LOBE_SIZE_MAX, dtype=dt) for _ in range(3)]
    ns = [ctypes.c_int(0) for _ in range(3)]
    p = ctypes.POINTER(ct)
    fn(sigma, bufs[0].ctypes.data_as(p), ctypes.byref(ns[0]),
       bufs[1].ctypes.data_as(p), ctypes.byref(ns[1]),
       bufs[2].ctypes.data_as(p), ctypes.byref(ns[2]))
    return tuple(b[: n.value].copy() for b, n in zip(bufs, ns))


def fir2(lobeH, lobeV, img: np.ndarray, out: np.ndarray, shift: bool = True,
         precision: str = "f32", nthreads: int = 1) -> np.ndarray:
    """FIRFilter2::initialize(lobeH, lobeV).convolve(img, out, shift)
    (FIRFilter.h:245-261, 298-377).  Writes ``out`` in place (only the valid
    region) and returns it.  ``precision`` "f32" is the faithful restatement,
    "f64" the double-precision one."""
    assert img.ndim == 2 and out.ndim == 2
    if img.strides[1] != img.itemsize or out.strides[1] != out.itemsize:
        raise ValueError("rows must be contiguous")
    nrow, ncol = img.shape
    if out.shape[0] < nrow or out.shape[1] < ncol:
        raise ValueError("output must be at least as large as the input")
    if precision == "f32":
        ct, dt, fn = ctypes.c_float, np.float32, lib().orc_fir2_f32
    else:
        ct, dt, fn = ctypes.c_double, np.float64, lib().orc_fir2_f64
    lh = np.ascontiguousarray(lobeH, dtype=dt)
    lv = np.ascontiguousarray(lobeV, dtype=dt)
    p = ctypes.POINTER(ct)
    rc = fn(lh.ctypes.data_as(p), lh.size, lv.ctypes.data_as(p), lv.size,
            img.ctypes.data, _dt(img), img.strides[0] // img.itemsize, nrow, ncol,
            out.ctypes.data, _dt(out), out.strides[0] // out.itemsize, int(shift), int(nthreads))
    if rc == 2:
        raise LobeTooLarge("FIRFilter2<T, BLOCK_TRAITS>::initialize: too large lobe size!")
    if rc == 3:
        raise UnsupportedLobe(f"FIRFilter2::convolve: unsupported lobe size [{lh.size}, {lv.size}]")
    if rc != 0:
        raise OracleError(f"oracle returned {rc}")
    return out


def fir2_taps(tapsH, tapsV, img: np.ndarray, out: np.ndarray, shift: bool = True,
              precision: str = "f32", nthreads: int = 1) -> np.ndarray:
    """General-taps separable filter (extension; FIRFilter2's window
    convention with odd tap counts up to 33, tap i weighting in[.][x+i])."""
    assert img.ndim == 2 and out.ndim == 2
    nrow, ncol = img.shape
    if out.shape[0] < nrow or out.shape[1] < ncol:
        raise ValueError("output must be at least as large as the input")
    if precision == "f32":
        ct, dt, fn = ctypes.c_float, np.float32, lib().orc_fir2_taps_f32
    else:
        ct, dt, fn = ctypes.c_double, np.float64, lib().orc_fir2_taps_f64
    th = np.ascontiguousarray(tapsH, dtype=dt)
    tv = np.ascontiguousarray(tapsV, dtype=dt)
    p = ctypes.POINTER(ct)
    rc = fn(th.ctypes.data_as(p), th.size, tv.ctypes.data_as(p), tv.size,
            img.ctypes.data, _dt(img), img.strides[0] // img.itemsize, nrow, ncol,
            out.ctypes.data, _dt(out), out.strides[0] // out.itemsize, int(shift), int(nthreads))
    if rc == 3:
        raise UnsupportedLobe("general taps: counts must be odd, 1 to 33")
    if rc != 0:
        raise OracleError(f"oracle returned {rc}")
    return out


def gauss(method: str, sigma: float, img: np.ndarray, out: np.ndarray, shift: bool = True,
          precision: str = "f32", nthreads: int = 1, lobes=None, rule: str = "reference") -> np.ndarray:
    """FIRGaussianConvolver2(sigma).<method>(img, out, shift)
    (FIRGaussianConvolver.h:178-242).  ``lobes`` overrides the lobes
    ``gauss_lobes(sigma, precision, rule)`` would give."""
    if lobes is None:
        lobes = gauss_lobes(sigma, precision, rule)
    h, v = METHODS[method]
    return fir2(lobes[h], lobes[v], img, out, shift, precision, nthreads)


def role_taps(lobe, symmetric: bool, precision: str = "f32") -> np.ndarray:
    """The 2r+1 taps of a Gaussian lobe whose parity comes from its role, not
    its length: lobe 0 and lobe 2 symmetric (r + 1 entries, centre last),
    lobe 1 antisymmetric (r entries, the left / upper side weighted +c[k],
    the other side -c[k], zero centre), lobe[0] outermost as
    FIRFilter.h:66-176 orders them."""
    lobe = np.asarray(lobe, np.float64)
    r = lobe.size - 1 if symmetric else lobe.size
    k = np.zeros(2 * r + 1)
    for i in range(1, r + 1):
        c = lobe[r - i]
        k[r - i] = c
        k[r + i] = c if symmetric else -c
    if symmetric:
        k[r] = lobe[r]
    return k.astype(np.float32 if precision == "f32" else np.float64)


def gauss_general(method: str, lobes, img: np.ndarray, out: np.ndarray, shift: bool = True,
                  precision: str = "f32", nthreads: int = 1) -> np.ndarray:
    """One of the six methods (FIRGaussianConvolver.h:178-242, the (H, V)
    lobe pairs of METHODS) for lobes of any length, e.g. those of the
    ceil(4 sigma) rule, whose symmetric lobes have an even number of entries
    when r is odd (FIRFilter2 would take such a lobe for an antisymmetric
    one).  Computed as general taps (``fir2_taps``), left tap first."""
    h, v = METHODS[method]
    return fir2_taps(role_taps(lobes[h], h != 1, precision), role_taps(lobes[v], v != 1, precision),
                     img, out, shift, precision, nthreads)


def subsample(img: np.ndarray, out: np.ndarray, op) -> np.ndarray:
    """cu::subsample(img, out, op) (algorithm.h:183-242): ``out[y][x]`` = the
    operator at ``img[2y][2x]`` for y < nrow/2, x < ncol/2, 0 where its
    stencil leaves the image (functional.h:58-256).  ``op`` is an index or a
    name of SUBSAMPLE_OPS.  Writes ``out`` in place and returns it."""
    code = SUBSAMPLE_OPS.index(op) if isinstance(op, str) else int(op)
    assert img.ndim == 2 and out.ndim == 2
    if img.strides[1] != img.itemsize or out.strides[1] != out.itemsize:
        raise ValueError("rows must be contiguous")
    nrow, ncol = img.shape
    if out.shape[0] < nrow // 2 or out.shape[1] < ncol // 2:
        raise ValueError("output must hold nrow/2 x ncol/2")
    rc = lib().orc_subsample_f32(img.ctypes.data, _dt(img), img.strides[0] // img.itemsize, nrow, ncol,
                                 out.ctypes.data, _dt(out), out.strides[0] // out.itemsize, code)
    if rc != 0:
        raise OracleError(f"oracle returned {rc}")
    return out


ICIA_MODELS = {"affine": 6, "projective": 8}


def icia_moment(edgeH: np.ndarray, edgeV: np.ndarray, model: str = "affine"):
    """ICIA's error moment (ICIA.h:61-112, 352-369): the sum over every pixel
    (row v, column u) of the full-size float gradient images of the upper
    triangle, row by row, of a a^T with a = image_derivative0(s u, s v, eH,
    eV) of the affine (vec.h:1418-1422, 6 entries) or projective
    (vec.h:1285-1290, 8 entries) map, s = 1 / max(nrow, ncol).  Returns
    (moment, abs_moment) in float64: the sums and the sums of the terms'
    magnitudes (the scale a summation's rounding error is measured on)."""
    D = ICIA_MODELS[model]
    eH = np.ascontiguousarray(edgeH, np.float32)
    eV = np.ascontiguousarray(edgeV, np.float32)
    assert eH.shape == eV.shape and eH.ndim == 2
    nm = D * (D + 1) // 2
    out, aout = np.zeros(nm), np.zeros(nm)
    p32, p64 = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
    rc = lib().orc_icia_moment(eH.ctypes.data_as(p32), eV.ctypes.data_as(p32), eH.shape[1],
                               eH.shape[0], eH.shape[1], D, out.ctypes.data_as(p64), aout.ctypes.data_as(p64))
    if rc != 0:
        raise OracleError(f"oracle returned {rc}")
    return out, aout


def full_kernel(lobe) -> np.ndarray:
    """The 2r+1 tap kernel a lobe stands for, ordered left (top) to right
    (bottom) as FIRFilter.h:66-176 applies it: tap i of the window has
    weight lobe[i] for i < r, lobe[r] at the centre (odd length only) and
    +-lobe[2r-i] for i > r (minus for even length)."""
    lobe = np.asarray(lobe, dtype=np.float64)
    L = lobe.size
    r = L & ~1
    k = np.zeros(2 * r + 1)
    for i in range(r):
        k[i] = lobe[i]
        k[2 * r - i] = lobe[i] if (L & 1) else -lobe[i]
    if L & 1:
        k[r] = lobe[r]
    return k


def half_to_float(h: int) -> float:
    return lib().orc_half_to_float(h)


def float_to_half(f: float) -> int:
    return lib().orc_float_to_half(f)


def bf16_to_float(b: int) -> float:
    return lib().orc_bf16_to_float(b)


def float_to_bf16(f: float) -> int:
    return lib().orc_float_to_bf16(f)


def float_to_u8(f: float) -> int:
    return lib().orc_float_to_u8(f)
