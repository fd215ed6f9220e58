"""Non-finite and extreme-magnitude pixels.

In the reference every output is a sum over its (2 rV + 1) x (2 rH + 1)
window with the taps of its lobes (FIRFilter.h:66-176, 181-231), so a NaN
or an Inf reaches exactly the outputs whose window holds it (an
antisymmetric lobe has no centre tap and leaves its centre pixel out; a
zero tap inside the lobe still multiplies), with the sign the tap products
give, and NaN where +Inf and -Inf meet.  The GPU outputs must have the
oracle's NaN, +Inf and -Inf masks exactly and match its finite values to
the FP32 bar.  Strips whose magnitudes lie far outside [0, 255] (1e-30,
1e35) must hold the bar relative to their own scale."""
import numpy as np
import pytest

import oracle
from tests._util import REL_TOL, assert_f32_parity, assert_u8_parity, random_lobe, to_dev, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu


def _planted(seed, nrow, ncol, n=7):
    """An edge image with n NaN, n +Inf and n -Inf pixels, some of them
    clustered (so that +Inf and -Inf share windows) and some at the border."""
    img = synth.edge_image(seed, nrow, ncol)
    rng = np.random.default_rng(seed)
    pts = [(int(rng.integers(0, nrow)), int(rng.integers(0, ncol))) for _ in range(3 * n - 4)]
    pts += [(0, 0), (nrow - 1, ncol // 2), (nrow // 2, 3), (nrow // 2 + 2, 5)]     # border, cluster
    for i, (y, x) in enumerate(pts):
        img[y, x] = (np.nan, np.inf, -np.inf)[i % 3]
    return img


def _compare(got, ref, what):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    for name, f in (("NaN", np.isnan), ("+Inf", np.isposinf), ("-Inf", np.isneginf)):
        g, r = f(got), f(ref)
        if not (g == r).all():
            idx = np.argwhere(g != r)[:5].tolist()
            raise AssertionError(f"{what}: {name} mask differs at {(g != r).sum()} pixels, first {idx}")
    fin = np.isfinite(ref)
    assert_f32_parity(got[fin], ref[fin], what)


@pytest.mark.parametrize("sigma", [1.0, 3.0])          # lobe radii (4, 4, 4) and (16, 16, 16)
@pytest.mark.parametrize("shift", [True, False])
def test_gauss6_nonfinite_masks(dev, kernel_family, sigma, shift):
    import torch
    import tucuda_amd
    img = _planted(11, 96, 200)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    outs = {m: torch.full(img.shape, -5.0, device=dev) for m in tucuda_amd.METHOD_ORDER}
    g.convolve6(to_dev(img, dev), outs, shift)
    torch.cuda.synchronize()
    lobes = oracle.gauss_lobes(sigma)
    for m in tucuda_amd.METHOD_ORDER:
        ref = np.full(img.shape, -5.0, np.float32)
        oracle.gauss(m, sigma, img, ref, shift, lobes=lobes)
        _compare(to_host(outs[m]), ref, f"{kernel_family} sigma={sigma} {m} shift={shift}")


@pytest.mark.parametrize("sigma", [1.2, 1.5, 2.5])     # mixed radii in one fused launch
def test_gauss6_nonfinite_mixed_radii(dev, sigma):
    """Lobes of different radii share one launch; each output still takes
    only its own lobes' taps (matrix-core kernels)."""
    import torch
    import tucuda_amd
    img = _planted(12, 80, 150)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    outs = {m: torch.full(img.shape, -5.0, device=dev) for m in tucuda_amd.METHOD_ORDER}
    g.convolve6(to_dev(img, dev), outs)
    torch.cuda.synchronize()
    lobes = oracle.gauss_lobes(sigma)
    for m in tucuda_amd.METHOD_ORDER:
        ref = np.full(img.shape, -5.0, np.float32)
        oracle.gauss(m, sigma, img, ref, lobes=lobes)
        _compare(to_host(outs[m]), ref, f"sigma={sigma} {m}")


@pytest.mark.parametrize("nH,nV", [(17, 16), (8, 9), (16, 4), (3, 2)])
def test_fir2_nonfinite_masks(dev, kernel_family, nH, nV):
    """FIRFilter2 lobes, symmetric and antisymmetric (no centre tap), one
    with a zero tap inside the lobe."""
    import torch
    import tucuda_amd
    rng = np.random.default_rng(nH * 31 + nV)
    lobeH = random_lobe(rng, nH)                     # unit L1 gain, so the FP32 bar applies as is
    lobeV = random_lobe(rng, nV)
    lobeH[1] = 0.0                                   # a zero tap inside the lobe still multiplies
    img = _planted(13, 70, 130)
    f = tucuda_amd.FIRFilter2().initialize(lobeH, lobeV)
    ref = np.full(img.shape, 7, np.float32)
    oracle.fir2(lobeH, lobeV, img, ref)
    out = to_dev(np.full(img.shape, 7, np.float32), dev)
    f.convolve(to_dev(img, dev), out)
    out8 = to_dev(np.full(img.shape, 7, np.uint8), dev)
    f.convolve(to_dev(img, dev), out8)
    torch.cuda.synchronize()
    _compare(to_host(out), ref, f"{kernel_family} lobes {nH}/{nV}")
    # u_char stores: NaN -> 0, +Inf -> 255, -Inf -> 0 (saturate, then truncate)
    ref8 = np.full(img.shape, 7, np.uint8)
    oracle.fir2(lobeH, lobeV, img, ref8)
    got8 = to_host(out8)
    nonfin = ~np.isfinite(ref)
    assert (got8[nonfin] == ref8[nonfin]).all()
    assert_u8_parity(got8[~nonfin], ref8[~nonfin], ref[~nonfin], f"u8 lobes {nH}/{nV}")


@pytest.mark.parametrize("nH,nV", [(17, 9), (5, 11), (33, 3)])
def test_general_taps_nonfinite_masks(dev, nH, nV):
    """General taps: every tap of the window multiplies, zeros included,
    none beyond it (matrix-core kernels)."""
    import torch
    import tucuda_amd
    rng = np.random.default_rng(nH + 100 * nV)
    th = rng.standard_normal(nH)
    tv = rng.standard_normal(nV)
    th, tv = (th / np.abs(th).sum()).astype(np.float32), (tv / np.abs(tv).sum()).astype(np.float32)
    th[nH // 2 - 1] = 0.0
    img = _planted(14, 75, 140)
    f = tucuda_amd.SeparableFilter2(th, tv)
    out = torch.full(img.shape, -5.0, device=dev)
    f.convolve(to_dev(img, dev), out)
    torch.cuda.synchronize()
    ref = np.full(img.shape, -5.0, np.float32)
    oracle.fir2_taps(th, tv, img, ref)
    _compare(to_host(out), ref, f"taps {nH}x{nV}")


@pytest.mark.parametrize("scale", [1e-30, 1e-36, 1e35])
def test_gauss6_extreme_magnitudes(dev, kernel_family, scale):
    """An image scaled far outside [0, 255] (every strip beyond the reach
    of the matrix-core kernel's power-of-two scale): every output within
    4e-5 of the largest output magnitude, the FP32 bar's relative part taken
    normwise."""
    import torch
    import tucuda_amd
    sigma = 2.0
    img = (synth.edge_image(15, 64, 256).astype(np.float64) * scale).astype(np.float32)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    outs = {m: torch.full(img.shape, np.nan, device=dev) for m in tucuda_amd.METHOD_ORDER}
    g.convolve6(to_dev(img, dev), outs)
    torch.cuda.synchronize()
    lobes = oracle.gauss_lobes(sigma, "f64")
    for m in tucuda_amd.METHOD_ORDER:
        ref = np.full(img.shape, np.nan, np.float32)
        oracle.gauss(m, sigma, img, ref, precision="f64", lobes=lobes)
        got = to_host(outs[m]).astype(np.float64)
        ref = ref.astype(np.float64)
        assert (np.isnan(got) == np.isnan(ref)).all(), m
        ok = ~np.isnan(ref)
        peak = np.abs(ref[ok]).max()
        err = np.abs(got[ok] - ref[ok]).max()
        assert err <= 4 * REL_TOL * peak, (kernel_family, scale, m, err, peak)
