"""Precision mode: FP16 / BF16 storage with FP32 accumulation, each output's
error stated against the FP64 oracle, and the general-taps extension.

Workloads: the 4096x4096 u_char image (seed 2) filtered by FIRFilter2 with
N(0,1) lobes normalised to unit L1 norm of the full kernel -- symmetric
9 / 5 (17 x 9 taps), antisymmetric 8 / 4, and non-symmetric 17 x 9 general
taps -- into float32 and float16; and the six Gaussian outputs with BF16
and with FP32 storage.  Errors are normwise: max|got - ref| / max|ref| and
rms(got - ref) / rms(ref), with ref the FP64 oracle on the exact input."""
import json
import os

import numpy as np
import pytest

import oracle
from tests._util import assert_f32_parity, bf16_bits_to_f32, to_dev, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

RESULTS = {}


def _errors(got: np.ndarray, ref: np.ndarray):
    d = got.astype(np.float64) - ref
    return {"max_rel": float(np.abs(d).max() / np.abs(ref).max()),
            "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref)))}


def _record(name, errs):
    """Keeps the error; with TUCU_RESULTS_DIR set, also writes every error so
    far to precision_errors.json there."""
    RESULTS[name] = errs
    out = os.environ.get("TUCU_RESULTS_DIR")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "precision_errors.json"), "w") as f:
            json.dump(RESULTS, f, indent=1)


def _lobes(seed, nH, nV):
    z = synth.normal(seed, nH + nV)
    lh, lv = z[:nH], z[nH:]
    lh = lh / np.abs(oracle.full_kernel(lh)).sum()
    lv = lv / np.abs(oracle.full_kernel(lv)).sum()
    return lh.astype(np.float32), lv.astype(np.float32)


@pytest.fixture(scope="module")
def img4k():
    return synth.u8_image(2, 4096, 4096)


@pytest.mark.parametrize("nH,nV", [(9, 5), (8, 4)])
def test_fir_4096_fp32_and_fp16(dev, img4k, nH, nV):
    import torch
    import tucuda_amd
    lh, lv = _lobes(2, nH, nV)
    r_v, r_h = nV & ~1, nH & ~1
    out64 = np.zeros(img4k.shape, np.float32)
    oracle.fir2(lh, lv, img4k, out64, True, "f64", nthreads=16)
    ref32 = np.zeros(img4k.shape, np.float32)
    oracle.fir2(lh, lv, img4k, ref32, True, "f32", nthreads=16)
    f = tucuda_amd.FIRFilter2().initialize(lh, lv)
    d_in = to_dev(img4k, dev)
    sl = (slice(r_v, 4096 - r_v), slice(r_h, 4096 - r_h))
    for dt, bound in ((torch.float32, 1e-5), (torch.float16, 1e-3)):
        d_out = torch.zeros(img4k.shape, dtype=dt, device=dev)
        f.convolve(d_in, d_out)
        got = to_host(d_out).astype(np.float32)
        if dt == torch.float32:
            assert_f32_parity(got, ref32, f"4096 {nH}/{nV} fp32")
        errs = _errors(got[sl], out64[sl].astype(np.float64))
        _record(f"fir_4096_u8_{nH}x{nV}_{str(dt).split('.')[-1]}", errs)
        assert errs["max_rel"] <= bound, errs


def test_general_taps_4096(dev, img4k):
    """Non-symmetric 17-tap horizontal, 9-tap vertical kernels (i.i.d.
    N(0,1), unit L1 norm) through the general-taps extension."""
    import torch
    import tucuda_amd
    z = synth.normal(22, 26)
    th, tv = z[:17], z[17:]
    th, tv = (th / np.abs(th).sum()).astype(np.float32), (tv / np.abs(tv).sum()).astype(np.float32)
    ref64 = np.zeros(img4k.shape, np.float32)
    oracle.fir2_taps(th, tv, img4k, ref64, True, "f64", nthreads=16)
    ref32 = np.zeros(img4k.shape, np.float32)
    oracle.fir2_taps(th, tv, img4k, ref32, True, "f32", nthreads=16)
    f = tucuda_amd.SeparableFilter2(th, tv)
    d_in = to_dev(img4k, dev)
    sl = (slice(4, 4092), slice(8, 4088))
    for dt, bound in ((torch.float32, 1e-5), (torch.float16, 1e-3)):
        d_out = torch.zeros(img4k.shape, dtype=dt, device=dev)
        f.convolve(d_in, d_out)
        got = to_host(d_out).astype(np.float32)
        if dt == torch.float32:
            assert_f32_parity(got, ref32, "4096 general taps fp32")
            assert (got[:4] == 0).all() and (got[:, :8] == 0).all()
        errs = _errors(got[sl], ref64[sl].astype(np.float64))
        _record(f"taps_4096_u8_17x9_{str(dt).split('.')[-1]}", errs)
        assert errs["max_rel"] <= bound, errs


@pytest.mark.parametrize("sigma", [0.8, 2.0, 3.0, 4.0])
def test_bf16_storage_gauss6(dev, sigma):
    """BF16 in, BF16 out, FP32 accumulation, against FP64 two ways:
    on the stored (BF16-rounded) input -- the error the filter adds, bound
    1e-2 normwise max -- and end to end on the original float image, where
    rounding the input to BF16 (8-bit mantissa) dominates the first
    derivatives at small sigma; bound 1e-2 normwise RMS, max recorded."""
    import torch
    import tucuda_amd
    img = synth.edge_image(4, 1080, 1920)
    lobes64 = oracle.gauss_lobes(sigma, "f64")
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    d_in = torch.from_numpy(img).to(dev).to(torch.bfloat16)
    stored = bf16_bits_to_f32(to_host(d_in))
    outs = {m: torch.zeros(img.shape, dtype=torch.bfloat16, device=dev) for m in oracle.METHOD_ORDER}
    g.convolve6(d_in, outs)
    for m in oracle.METHOD_ORDER:
        h, v = oracle.METHODS[m]
        got = bf16_bits_to_f32(to_host(outs[m]))
        for name, src in (("stored", stored), ("end_to_end", img)):
            ref = np.full(img.shape, np.nan, np.float32)
            oracle.fir2(lobes64[h], lobes64[v], src, ref, True, "f64", nthreads=16)
            valid = ~np.isnan(ref)
            errs = _errors(got[valid], ref[valid].astype(np.float64))
            _record(f"gauss6_1080p_bf16_sigma{sigma}_{m}_{name}", errs)
            if name == "stored":
                assert errs["max_rel"] <= 1e-2, (m, name, errs)
            else:
                assert errs["rms_rel"] <= 1e-2, (m, name, errs)


@pytest.mark.parametrize("sigma", [0.8, 2.0, 3.0, 4.0])
def test_f32_gauss6_against_fp64(dev, sigma):
    """FP32 storage through the default six-output kernel (matrix cores,
    f16 hi / lo operand splits, FP32 accumulation) against the FP64 oracle:
    normwise max error <= 1e-5, the FP32 parity bar, and recorded."""
    import torch
    import tucuda_amd
    img = synth.edge_image(4, 1080, 1920)
    lobes64 = oracle.gauss_lobes(sigma, "f64")
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    outs = {m: torch.zeros(img.shape, dtype=torch.float32, device=dev) for m in oracle.METHOD_ORDER}
    g.convolve6(to_dev(img, dev), outs)
    for m in oracle.METHOD_ORDER:
        h, v = oracle.METHODS[m]
        ref = np.full(img.shape, np.nan, np.float32)
        oracle.fir2(lobes64[h], lobes64[v], img, ref, True, "f64", nthreads=16)
        valid = ~np.isnan(ref)
        errs = _errors(to_host(outs[m])[valid], ref[valid].astype(np.float64))
        _record(f"gauss6_1080p_f32_sigma{sigma}_{m}", errs)
        assert errs["max_rel"] <= 1e-5, (m, errs)


def test_fp16_storage_matches_oracle_storage_rounding(dev):
    """With FP16 output storage the kernel and the oracle round FP32 results
    that agree to the FP32 parity bar: equal within one FP16 step plus that
    bar (near zero the FP16 subnormal step is finer than FP32 rounding of a
    sum of terms of size ~20)."""
    import torch
    import tucuda_amd
    img = synth.u8_image(5, 300, 520)
    lh, lv = _lobes(5, 17, 9)
    ref = np.zeros(img.shape, np.float16)
    oracle.fir2(lh, lv, img, ref)
    d_out = torch.zeros(img.shape, dtype=torch.float16, device=dev)
    tucuda_amd.FIRFilter2().initialize(lh, lv).convolve(to_dev(img, dev), d_out)
    got = to_host(d_out)
    ulp = np.spacing(np.abs(ref).astype(np.float16)).astype(np.float32)
    tol = ulp + np.maximum(1e-5 * np.abs(ref.astype(np.float32)), 5e-6 * 255)
    assert (np.abs(got.astype(np.float32) - ref.astype(np.float32)) <= tol).all()
