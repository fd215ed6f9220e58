"""Python mirror of TU::cu::FIRFilter2, TU::cu::FIRGaussianConvolver2 and
TU::cu::subsample.

Same method names, argument meanings and error behaviour as the C++
classes and function (TU/cu/FIRFilter.h:18-52, TU/cu/FIRGaussianConvolver.h:56-90,
TU/cu/algorithm.h:183-242);
images are 2-D device tensors (rows contiguous, any row stride) standing in
for ``cu::Array2<T>`` row ranges.  All work is enqueued on torch's current
HIP stream through the C ABI; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes
from typing import Mapping, Optional, Sequence, Union

import numpy as np
import torch

from . import _capi
from ._capi import TucuImage, check

_DTYPES = {
    torch.uint8: _capi.TUCU_U8,
    torch.float32: _capi.TUCU_F32,
    torch.float16: _capi.TUCU_F16,
    torch.bfloat16: _capi.TUCU_BF16,
}

METHOD_ORDER = ("smooth", "diffH", "diffV", "diffHH", "diffHV", "diffVV")
METHOD_BITS = {m: 1 << i for i, m in enumerate(METHOD_ORDER)}
# (horizontal lobe, vertical lobe): FIRGaussianConvolver.h:181, 193, 205, 217, 229, 241
METHOD_LOBES = {"smooth": (0, 0), "diffH": (1, 0), "diffV": (0, 1),
                "diffHH": (2, 0), "diffHV": (1, 1), "diffVV": (0, 2)}


def image_desc(t: torch.Tensor, with_extent: bool = True) -> TucuImage:
    """The C descriptor of a 2-D device tensor."""
    if not isinstance(t, torch.Tensor) or t.dim() != 2:
        raise ValueError("images are 2-D tensors")
    if t.dtype not in _DTYPES:
        raise TypeError(f"unsupported element type {t.dtype}")
    if not t.is_cuda:
        raise ValueError("images must live in device memory")
    if t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError("image rows must be contiguous")
    stride = t.stride(0) if t.shape[0] > 1 else max(t.shape[1], 1)
    return TucuImage(t.data_ptr(), stride,
                     t.shape[0] if with_extent else 0,
                     t.shape[1] if with_extent else 0,
                     _DTYPES[t.dtype], 0)


def array2(nrow: int, ncol: int, dtype=torch.float32, device="cuda") -> torch.Tensor:
    """A zero-filled device image laid out like ``cu::Array2<T>``: rows
    padded to a multiple of 256 bytes (allocator.h:68, Array++.h:142),
    returned as the ``nrow x ncol`` view.  Rows that start on cache-line
    boundaries let the kernels write whole lines; with packed rows of an
    arbitrary width, the partial lines at tile edges cost the write stream
    much of its rate."""
    if dtype not in _DTYPES:
        raise TypeError(f"unsupported element type {dtype}")
    es = torch.empty((), dtype=dtype).element_size()
    stride = int(_capi.load().tucu_row_stride(int(ncol), es))
    return torch.zeros((int(nrow), max(stride, 1)), dtype=dtype, device=device)[:, :int(ncol)]


def _stream_handle(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _f32(lobe) -> np.ndarray:
    if isinstance(lobe, torch.Tensor):
        lobe = lobe.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(lobe, dtype=np.float32).reshape(-1))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class FIRFilter2:
    """Separable 2-D FIR filter with symmetric (odd length) or antisymmetric
    (even length) half-kernels -- TU::cu::FIRFilter2 (FIRFilter.h:18-52)."""

    LobeSizeMax = 17                          # FIRFilter.h:26

    def __init__(self):
        self._lobeH = np.zeros(0, np.float32)
        self._lobeV = np.zeros(0, np.float32)

    def initialize(self, lobeH, lobeV) -> "FIRFilter2":
        """FIRFilter.h:245-261: LobeTooLarge (a RuntimeError) when a lobe is
        longer than 17; the lengths are validated against the supported set
        at convolve time."""
        lh, lv = _f32(lobeH), _f32(lobeV)
        check(_capi.load().tucu_fir2_check(lh.size, lv.size))
        self._lobeH, self._lobeV = lh, lv
        return self

    def convolve(self, inp: torch.Tensor, out: torch.Tensor, shift: bool = True) -> None:
        """FIRFilter.h:298-377: writes the (nrow-2*offsetV) x (ncol-2*offsetH)
        valid region of ``out`` (at (offsetV, offsetH) if ``shift``), leaves
        the rest untouched; images too small for the kernel are a no-op."""
        src, dst = image_desc(inp), image_desc(out)
        if inp.device != out.device:
            raise ValueError("input and output must be on the same device")
        check(_capi.load().tucu_fir2_convolve(
            _fptr(self._lobeH), self._lobeH.size, _fptr(self._lobeV), self._lobeV.size,
            ctypes.byref(src), ctypes.byref(dst), int(bool(shift)), _stream_handle(inp)))

    def outSizeH(self, inSize: int) -> int:     # FIRFilter.h:37
        return inSize - 2 * self.offsetH()

    def outSizeV(self, inSize: int) -> int:     # FIRFilter.h:38
        return inSize - 2 * self.offsetV()

    def offsetH(self) -> int:                   # FIRFilter.h:39
        return self._lobeH.size & ~1

    def offsetV(self) -> int:                   # FIRFilter.h:40
        return self._lobeV.size & ~1


class SeparableFilter2:
    """Extension: separable filter with general (non-symmetric) taps, the
    window convention of FIRFilter2 (tap i weights in[y][x+i] / in[y+i][x]).
    Odd tap counts 1..33; output placement as FIRFilter2 with r = (n-1)/2."""

    def __init__(self, tapsH=None, tapsV=None):
        self._th = _f32(tapsH if tapsH is not None else [1.0])
        self._tv = _f32(tapsV if tapsV is not None else [1.0])

    def initialize(self, tapsH, tapsV) -> "SeparableFilter2":
        self._th, self._tv = _f32(tapsH), _f32(tapsV)
        return self

    def offsetH(self) -> int:
        return self._th.size // 2

    def offsetV(self) -> int:
        return self._tv.size // 2

    def convolve(self, inp: torch.Tensor, out: torch.Tensor, shift: bool = True) -> None:
        if inp.device != out.device:
            raise ValueError("input and output must be on the same device")
        src, dst = image_desc(inp), image_desc(out)
        check(_capi.load().tucu_fir2_convolve_taps(
            _fptr(self._th), self._th.size, _fptr(self._tv), self._tv.size,
            ctypes.byref(src), ctypes.byref(dst), int(bool(shift)), _stream_handle(inp)))


SUBSAMPLE_OPS = _capi.SUBSAMPLE_OPS


def subsample(inp: torch.Tensor, out: torch.Tensor, op="identity1x1") -> None:
    """cu::subsample(in, ie, out, op) (algorithm.h:183-242): ``out[y][x]`` =
    the operator (a name of SUBSAMPLE_OPS or its index, functional.h:58-256)
    at ``inp[2y][2x]`` for y < nrow/2, x < ncol/2, 0 where its stencil would
    leave the image.  Images with fewer than 2 rows or columns are a no-op.
    Operators other than identity1x1 need a floating-point input."""
    code = SUBSAMPLE_OPS.index(op) if isinstance(op, str) else int(op)
    if inp.device != out.device:
        raise ValueError("input and output must be on the same device")
    src, dst = image_desc(inp), image_desc(out)
    check(_capi.load().tucu_subsample(ctypes.byref(src), ctypes.byref(dst), code, _stream_handle(inp)))


def gauss_lobes(sigma: float, rule: int = _capi.TUCU_RULE_REFERENCE):
    """FIRGaussianConvolver2::initialize(sigma, lobe0, lobe1, lobe2)
    (FIRGaussianConvolver.h:123-170), computed by the library on the host."""
    bufs = [np.zeros(17, np.float32) for _ in range(3)]
    ns = [ctypes.c_int(0) for _ in range(3)]
    check(_capi.load().tucu_gauss2_lobes(
        ctypes.c_float(sigma), rule,
        _fptr(bufs[0]), ctypes.byref(ns[0]), _fptr(bufs[1]), ctypes.byref(ns[1]),
        _fptr(bufs[2]), ctypes.byref(ns[2])))
    return tuple(b[: n.value].copy() for b, n in zip(bufs, ns))


def lobe_size(lobe17, even: bool) -> int:
    """detail::lobeSize (FIRGaussianConvolver.cu:52-98)."""
    a = _f32(lobe17)
    if a.size != 17:
        raise ValueError("lobeSize takes 17 samples")
    return _capi.load().tucu_lobe_size(_fptr(a), int(bool(even)))


OutputSpec = Union[Mapping[str, torch.Tensor], Sequence[Optional[torch.Tensor]]]


class FIRGaussianConvolver2(FIRFilter2):
    """Gaussian smoothing and derivatives -- TU::cu::FIRGaussianConvolver2
    (FIRGaussianConvolver.h:56-90)."""

    def __init__(self, sigma: float = 1.0, rule: int = _capi.TUCU_RULE_REFERENCE):
        super().__init__()
        self._rule = rule
        self.initialize(sigma)

    def sigma(self) -> float:
        return self._sigma

    def initialize(self, sigma: float) -> "FIRGaussianConvolver2":      # :114-121
        self._sigma = float(np.float32(sigma))
        self._lobe0, self._lobe1, self._lobe2 = gauss_lobes(self._sigma, self._rule)
        return self

    @staticmethod
    def lobes(sigma: float, rule: int = _capi.TUCU_RULE_REFERENCE):
        """The static initialize(sigma, lobe0, lobe1, lobe2) (:123-170)."""
        return gauss_lobes(sigma, rule)

    def _lobes(self):
        return (self._lobe0, self._lobe1, self._lobe2)

    def _method(self, name: str, inp, out, shift):
        if self._rule != _capi.TUCU_RULE_REFERENCE:
            self.convolve6(inp, {name: out}, shift)
            return
        h, v = METHOD_LOBES[name]
        lobes = self._lobes()
        FIRFilter2.initialize(self, lobes[h], lobes[v]).convolve(inp, out, shift)

    def smooth(self, inp, out, shift: bool = True):   # :178-182
        self._method("smooth", inp, out, shift)

    def diffH(self, inp, out, shift: bool = True):    # :190-194
        self._method("diffH", inp, out, shift)

    def diffV(self, inp, out, shift: bool = True):    # :202-206
        self._method("diffV", inp, out, shift)

    def diffHH(self, inp, out, shift: bool = True):   # :214-218
        self._method("diffHH", inp, out, shift)

    def diffHV(self, inp, out, shift: bool = True):   # :226-230
        self._method("diffHV", inp, out, shift)

    def diffVV(self, inp, out, shift: bool = True):   # :238-242
        self._method("diffVV", inp, out, shift)

    def smoothSubsample(self, inp: torch.Tensor, out: torch.Tensor, shift: bool = True) -> None:
        """Extension: the pyramid step in one launch.  ``out`` (at least
        nrow/2 x ncol/2) receives what ``smooth`` into a zero-filled image
        followed by ``subsample(..., "identity1x1")`` gives, without the
        full-size intermediate."""
        if inp.device != out.device:
            raise ValueError("input and output must be on the same device")
        src, dst = image_desc(inp), image_desc(out)
        check(_capi.load().tucu_gauss2_pyramid(_fptr(self._lobe0), self._lobe0.size, ctypes.byref(src),
                                               ctypes.byref(dst), int(bool(shift)), _stream_handle(inp)))

    def convolve6(self, inp: torch.Tensor, outs: OutputSpec, shift: bool = True) -> None:
        """Extension: every requested method from one fused launch.  ``outs``
        maps method names to output tensors, or lists up to six tensors in
        method order (None = not wanted).  Each output gets exactly what the
        corresponding single method would write."""
        if isinstance(outs, Mapping):
            seq = [outs.get(m) for m in METHOD_ORDER]
            unknown = set(outs) - set(METHOD_ORDER)
            if unknown:
                raise KeyError(f"unknown methods {sorted(unknown)}")
        else:
            seq = list(outs) + [None] * (6 - len(outs))
        descs = (TucuImage * 6)()
        mask = 0
        for k, t in enumerate(seq[:6]):
            if t is None:
                continue
            if t.device != inp.device:
                raise ValueError("input and outputs must be on the same device")
            descs[k] = image_desc(t)
            mask |= 1 << k
        src = image_desc(inp)
        l0, l1, l2 = self._lobes()
        check(_capi.load().tucu_gauss2_convolve(
            _fptr(l0), l0.size, _fptr(l1), l1.size, _fptr(l2), l2.size,
            ctypes.byref(src), descs, mask, int(bool(shift)), _stream_handle(inp)))


ICIA_MODELS = {"affine": _capi.TUCU_ICIA_AFFINE, "projective": _capi.TUCU_ICIA_PROJECTIVE}


def icia_moment(sigma_or_lobes, inp: torch.Tensor, model: str = "affine",
                rule: int = _capi.TUCU_RULE_REFERENCE) -> torch.Tensor:
    """ICIA's error moment (ICIA.h:346-369): diffH and diffV of ``inp`` from
    one fused pass that never writes them, and the sum over every pixel of
    the upper triangle of a a^T, a = image_derivative0(s u, s v, eH, eV) of
    the affine (21 moments) or projective (36) map.  Returns a float32
    device tensor of the moments, enqueued on torch's current stream."""
    if isinstance(sigma_or_lobes, (int, float)):
        l0, l1, _ = gauss_lobes(float(sigma_or_lobes), rule)
    else:
        l0, l1 = _f32(sigma_or_lobes[0]), _f32(sigma_or_lobes[1])
    m = ICIA_MODELS[model]
    src = image_desc(inp)
    need = ctypes.c_size_t(0)
    lib = _capi.load()
    check(lib.tucu_icia_moment(_fptr(l0), l0.size, _fptr(l1), l1.size, ctypes.byref(src), m, None, None,
                               ctypes.byref(need), _stream_handle(inp)))
    ws = torch.empty(int(need.value), dtype=torch.uint8, device=inp.device)
    out = torch.empty(m * (m + 1) // 2, dtype=torch.float32, device=inp.device)
    check(lib.tucu_icia_moment(_fptr(l0), l0.size, _fptr(l1), l1.size, ctypes.byref(src), m,
                               ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                               ctypes.byref(need), _stream_handle(inp)))
    return out


def batch_tile_rows():
    """Tile heights the batched kernels offer: the matrix-core kernels
    (default) take 16, 32 or 64 rows (workgroups of 4 x 1, 2 x 2 or 1 x 4
    strips of 16 x 64 pixels), the direct FP32 kernels (TUCU_GAUSS_PATH=valu)
    4, 8 or 16."""
    import os
    return (4, 8, 16) if os.environ.get("TUCU_GAUSS_PATH") == "valu" else (16, 32, 64)


class GaussBatch:
    """Batched mode (extension): many Gaussian jobs of uneven sizes and
    sigmas in one persistent launch.  ``jobs`` is a sequence of
    ``(input, outputs, lobes_or_sigma)`` or ``(input, outputs, lobes_or_sigma,
    shift)``; ``outputs`` maps method names to tensors (or lists six
    tensors / None); a float sigma takes lobes by ``rule``.  Each job gets
    exactly what FIRGaussianConvolver2.convolve6 would write for it.  The
    tensors must outlive the batch."""

    def __init__(self, jobs, rule: int = _capi.TUCU_RULE_REFERENCE):
        arr = (_capi.TucuGaussJob * max(1, len(jobs)))()
        keep = []
        self._device = None
        for i, job in enumerate(jobs):
            inp, outs, lob = job[0], job[1], job[2]
            shift = job[3] if len(job) > 3 else True
            if isinstance(lob, (int, float)):
                lob = gauss_lobes(float(lob), rule)
            lob = [_f32(x) for x in lob]
            keep.append(lob)
            if isinstance(outs, Mapping):
                seq = [outs.get(m) for m in METHOD_ORDER]
            else:
                seq = list(outs) + [None] * (6 - len(outs))
            j = arr[i]
            j.in_ = image_desc(inp)
            mask = 0
            for k, t in enumerate(seq[:6]):
                if t is not None:
                    j.out[k] = image_desc(t)
                    mask |= 1 << k
            j.lobe0, j.lobe1, j.lobe2 = (_fptr(x) for x in lob)
            j.n0, j.n1, j.n2 = (x.size for x in lob)
            j.mask, j.shift = mask, int(bool(shift))
            self._device = inp.device
        h = ctypes.c_void_p()
        check(_capi.load().tucu_batch_create(ctypes.byref(h), arr, len(jobs)))
        self._h = h

    def _stream(self):
        if self._device is None:
            return ctypes.c_void_p(None)
        return ctypes.c_void_p(torch.cuda.current_stream(self._device).cuda_stream)

    def launch(self) -> None:
        check(_capi.load().tucu_batch_launch(self._h, self._stream()))

    def set_tile_rows(self, rows: int) -> None:
        check(_capi.load().tucu_batch_set_tile_rows(self._h, int(rows)))

    def autotune(self, reps: int = 5):
        rows, us = ctypes.c_int(0), ctypes.c_float(0.0)
        check(_capi.load().tucu_batch_autotune(self._h, self._stream(), int(reps),
                                               ctypes.byref(rows), ctypes.byref(us)))
        return rows.value, us.value

    def info(self) -> dict:
        v = [ctypes.c_int(0) for _ in range(4)]
        check(_capi.load().tucu_batch_info(self._h, *(ctypes.byref(x) for x in v)))
        return dict(zip(("jobs", "tiles", "grid", "tile_rows"), (x.value for x in v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _capi.load().tucu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_cache_save(path: str) -> None:
    check(_capi.load().tucu_tune_cache_save(path.encode()))


def tune_cache_load(path: str) -> None:
    check(_capi.load().tucu_tune_cache_load(path.encode()))
