"""Helpers shared by the tests: tolerances, numpy <-> torch, oracle calls."""
from __future__ import annotations

import numpy as np

import oracle

# FP32 parity bar: |got - ref| <= max(1e-5 |ref|, 5e-6 * 255) on [0,255] data
# ("1e-5 relative, 5e-6 absolute on [0,1]-scaled data").
REL_TOL = 1e-5
ABS_TOL = 5e-6 * 255.0


def assert_f32_parity(got: np.ndarray, ref: np.ndarray, what: str = "") -> None:
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape
    same_nan = np.isnan(got) & np.isnan(ref)
    err = np.where(same_nan, 0.0, np.abs(got - ref))
    tol = np.maximum(REL_TOL * np.abs(ref), ABS_TOL)
    bad = ~(err <= tol)
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError(
            f"{what}: {bad.sum()} of {bad.size} elements out of tolerance; first at {idx.tolist()}: "
            f"got {[got[tuple(i)] for i in idx]} ref {[ref[tuple(i)] for i in idx]}")


def assert_u8_parity(got: np.ndarray, ref_u8: np.ndarray, ref_f: np.ndarray, what: str = "") -> None:
    """u_char outputs: equal to the oracle's, except that where the oracle's
    float result lies within the FP32 tolerance of an integer (a truncation
    edge) either neighbouring byte is accepted."""
    got = got.astype(np.int32)
    ref = ref_u8.astype(np.int32)
    f = ref_f.astype(np.float64)
    tol = np.maximum(REL_TOL * np.abs(f), ABS_TOL)
    near_edge = np.abs(f - np.round(f)) <= tol
    ok = (got == ref) | (near_edge & (np.abs(got - ref) <= 1))
    if not ok.all():
        idx = np.argwhere(~ok)[:5]
        raise AssertionError(
            f"{what}: {(~ok).sum()} u8 mismatches; first at {idx.tolist()}: got "
            f"{[got[tuple(i)] for i in idx]} ref {[ref[tuple(i)] for i in idx]} "
            f"float {[f[tuple(i)] for i in idx]}")


def to_dev(a: np.ndarray, dev):
    import torch
    if a.dtype == np.uint16:          # bf16 bit patterns
        return torch.from_numpy(a.view(np.int16).copy()).to(dev).view(torch.bfloat16)
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_host(t) -> np.ndarray:
    import torch
    if t.dtype == torch.bfloat16:
        return t.view(torch.int16).cpu().numpy().view(np.uint16)
    return t.cpu().numpy()


def bf16_bits_to_f32(a: np.ndarray) -> np.ndarray:
    return (a.astype(np.uint32) << 16).view(np.float32)


def sentinel(dtype) -> float:
    return 77 if dtype == np.uint8 else -12345.0


def random_lobe(rng: np.random.Generator, n: int) -> np.ndarray:
    """n i.i.d. N(0,1) lobe entries scaled so the full kernel has unit L1 norm."""
    lobe = rng.standard_normal(n)
    full = np.abs(oracle.full_kernel(lobe)).sum()
    return (lobe / full).astype(np.float32)


def full_taps(lobe, symmetric: bool) -> np.ndarray:
    """The 2r+1 taps of a lobe (outermost first; symmetric lobes end with the
    centre tap, antisymmetric ones weight the left side + and the right -)."""
    lobe = np.asarray(lobe, np.float64)
    r = lobe.size - 1 if symmetric else lobe.size
    k = np.zeros(2 * r + 1)
    for i in range(1, r + 1):
        c = lobe[r - i]
        k[r - i] = c
        k[r + i] = c if symmetric else -c
    if symmetric:
        k[r] = lobe[r]
    return k
