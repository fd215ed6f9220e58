"""Six-output Gaussian on every element-type pair and on unaligned views.

The tile grid starts its columns on a 128-byte boundary of the first output
and takes the check-free store path only where every output is aligned; views
whose rows start at odd offsets (input, outputs, and outputs misaligned with
each other) exercise the checked paths and the edge masking.  Reference: the
oracle on the same input with the same output element type (its own
conversion), so u8 / f16 / bf16 outputs may differ only by the rounding of a
result within the FP32 bar of a rounding edge."""
import numpy as np
import pytest

import oracle
from tests._util import ABS_TOL, REL_TOL, assert_f32_parity, assert_u8_parity, bf16_bits_to_f32, to_dev, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

TYPES = {"u8": np.uint8, "f32": np.float32, "f16": np.float16, "bf16": np.uint16}


def _input(kind, nrow, ncol):
    img = synth.edge_image(21, nrow, ncol)
    if kind == "u8":
        return np.clip(img, 0, 255).astype(np.uint8)
    if kind == "f16":
        return img.astype(np.float16)
    if kind == "bf16":
        return (img.view(np.uint32) >> 16).astype(np.uint16)    # truncation: exact bf16 values
    return img


def _as_f32(a, kind):
    if kind == "bf16":
        return bf16_bits_to_f32(a)
    return a.astype(np.float32)


def _check(got, ref, ref_f, kind, what):
    if kind == "f32":
        assert_f32_parity(got, ref, what)
    elif kind == "u8":
        assert_u8_parity(got, ref, ref_f, what)
    else:
        # one unit in the last place of the 16-bit result, where the FP32
        # result lies within the FP32 bar of a rounding edge
        g, r = _as_f32(got, kind).astype(np.float64), _as_f32(ref, kind).astype(np.float64)
        ulp = np.abs(r) * (2.0 ** -10 if kind == "f16" else 2.0 ** -7) + 1e-30
        tol = np.maximum(REL_TOL * np.abs(ref_f), ABS_TOL) + ulp
        bad = ~(np.abs(g - r) <= tol)
        assert not bad.any(), (what, int(bad.sum()), np.argwhere(bad)[:3].tolist())


@pytest.mark.parametrize("ik", ["u8", "f32", "f16", "bf16"])
@pytest.mark.parametrize("ok", ["u8", "f32", "f16", "bf16"])
def test_type_pairs_on_unaligned_views(dev, kernel_family, ik, ok):
    import torch
    import tucuda_amd
    sigma = 2.4                                # reference-rule lobes 17 / 16 / 17 (radius class 16)
    nrow, ncol = 75, 203
    src = _input(ik, nrow + 1, ncol + 3)      # the input view starts at (1, 3)
    d_src = to_dev(src, dev)[1:, 3:]
    img = src[1:, 3:]
    img_f = _as_f32(np.ascontiguousarray(img), ik)
    # outputs: views at different column offsets (misaligned with each other)
    sentinel = 7 if ok == "u8" else -7.0
    bigs = [np.full((nrow + 2, ncol + 8), sentinel, TYPES[ok]) if ok != "bf16" else
            np.full((nrow + 2, ncol + 8), 0xC0E0, np.uint16) for _ in range(6)]
    d_bigs = [to_dev(b, dev) for b in bigs]
    views = [d[2:, 1 + k:1 + k + ncol] for k, d in enumerate(d_bigs)]
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    g.convolve6(d_src, dict(zip(tucuda_amd.METHOD_ORDER, views)))
    torch.cuda.synchronize()
    lobes = oracle.gauss_lobes(sigma)
    for k, m in enumerate(tucuda_amd.METHOD_ORDER):
        got_big = to_host(d_bigs[k])
        got = got_big[2:, 1 + k:1 + k + ncol]
        ref = bigs[k][2:, 1 + k:1 + k + ncol].copy()
        ref_f = np.zeros(img.shape, np.float32)
        oracle.gauss(m, sigma, np.ascontiguousarray(img), ref, True, lobes=lobes)
        oracle.gauss(m, sigma, img_f, ref_f, True, lobes=lobes)
        _check(got, ref, ref_f, ok, f"{ik}->{ok} {m} ({kernel_family})")
        # nothing written outside the view
        outside = np.ones(got_big.shape, bool)
        outside[2:, 1 + k:1 + k + ncol] = False
        assert (got_big[outside] == bigs[k][outside]).all(), (ik, ok, m)


@pytest.mark.parametrize("dtype_name", ["uint8", "float32", "float16", "bfloat16"])
def test_array2_layout(dev, dtype_name):
    """tucuda_amd.array2: the cu::Array2 layout (rows padded to a multiple of
    256 bytes, zero-filled), a view the filters take like any other."""
    import torch
    import tucuda_amd
    dt = getattr(torch, dtype_name)
    t = tucuda_amd.array2(37, 101, dt, dev)
    es = t.element_size()
    assert tuple(t.shape) == (37, 101) and t.stride(1) == 1
    assert (t.stride(0) * es) % 256 == 0 and t.stride(0) * es >= 101 * es
    assert t.stride(0) * es - 101 * es < 256
    assert (t.data_ptr() % 256) == 0
    assert bool((t.float() == 0).all())
