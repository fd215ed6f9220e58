"""FIRFilter2 parity on the GPU: every supported (horizontal, vertical) lobe
length pair x {u_char, float} in and out (the instantiations of
FIRFilter.cu:49-68) x shift, against the oracle's faithful FP32
restatement on the same seeded inputs.  Outputs are pre-filled with a
sentinel on both sides, so the comparison also checks that nothing outside
the valid region is written (FIRFilter.h:37-40, 286-287)."""
import itertools

import numpy as np
import pytest

import oracle
from tests._util import (assert_f32_parity, assert_u8_parity, random_lobe, sentinel,
                         to_dev, to_host)
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

SIZES = oracle.SUPPORTED_LOBES
NP_TYPES = {"u8": np.uint8, "f32": np.float32}


def _input(kind: str, seed: int, nrow: int, ncol: int) -> np.ndarray:
    if kind == "u8":
        return synth.u8_image(seed, nrow, ncol)
    return synth.uniform_image(seed, nrow, ncol)


def _run(dev, lobeH, lobeV, img, out_np_dtype, shift, out_shape=None):
    import tucuda_amd
    nrow, ncol = img.shape
    shape = out_shape or (nrow, ncol)
    ref = np.full(shape, sentinel(out_np_dtype), dtype=out_np_dtype)
    oracle.fir2(lobeH, lobeV, img, ref, shift)
    reff = np.full(shape, sentinel(np.float32), dtype=np.float32)
    oracle.fir2(lobeH, lobeV, img, reff, shift)
    d_in = to_dev(img, dev)
    d_out = to_dev(np.full(shape, sentinel(out_np_dtype), dtype=out_np_dtype), dev)
    f = tucuda_amd.FIRFilter2().initialize(lobeH, lobeV)
    f.convolve(d_in, d_out, shift)
    got = to_host(d_out)
    return got, ref, reff


@pytest.mark.parametrize("nH,nV", list(itertools.product(SIZES, SIZES)))
def test_fir_all_lobe_pairs(dev, kernel_family, nH, nV):
    rng = np.random.default_rng(1000 + 17 * nH + nV)
    lobeH, lobeV = random_lobe(rng, nH), random_lobe(rng, nV)
    # 300 columns = more than one 256-wide strip at every radius; odd rows.
    for (ik, ok), shift in itertools.product(itertools.product(("u8", "f32"), ("u8", "f32")), (True, False)):
        img = _input(ik, nH * 31 + nV, 71, 300)
        got, ref, reff = _run(dev, lobeH, lobeV, img, NP_TYPES[ok], shift)
        what = f"nH={nH} nV={nV} {ik}->{ok} shift={shift}"
        if ok == "u8":
            assert_u8_parity(got, ref, reff, what)
        else:
            assert_f32_parity(got, ref, what)


@pytest.mark.parametrize("nrow,ncol", [(33, 33), (34, 35), (35, 34), (64, 257), (129, 481), (5, 700)])
def test_fir_ragged_sizes(dev, kernel_family, nrow, ncol):
    """Sizes just above the no-op threshold, odd sizes, strip boundaries."""
    rng = np.random.default_rng(nrow * 1000 + ncol)
    for nH, nV in [(17, 17), (16, 9), (3, 2), (9, 16), (5, 4)]:
        lobeH, lobeV = random_lobe(rng, nH), random_lobe(rng, nV)
        img = synth.uniform_image(nrow + ncol, nrow, ncol)
        got, ref, _ = _run(dev, lobeH, lobeV, img, np.float32, True)
        assert_f32_parity(got, ref, f"{nrow}x{ncol} nH={nH} nV={nV}")


def test_fir_too_small_is_noop(dev, kernel_family):
    """nrow <= 2 rV or ncol <= 2 rH: nothing is written (FIRFilter.h:304-312)."""
    lobe = random_lobe(np.random.default_rng(0), 17)
    for nrow, ncol in [(32, 100), (100, 32), (10, 10)]:
        img = synth.uniform_image(1, nrow, ncol)
        got, ref, _ = _run(dev, lobe, lobe, img, np.float32, True)
        assert (got == sentinel(np.float32)).all()
        assert (ref == sentinel(np.float32)).all()


def test_fir_errors(dev):
    import tucuda_amd
    f = tucuda_amd.FIRFilter2()
    with pytest.raises(tucuda_amd.LobeTooLarge):
        f.initialize(np.zeros(18), np.zeros(3))
    d_in = to_dev(synth.uniform_image(2, 64, 64), dev)
    d_out = to_dev(np.zeros((64, 64), np.float32), dev)
    f.initialize(np.ones(6) / 11, np.ones(3) / 5)
    with pytest.raises(tucuda_amd.UnsupportedLobe, match="horizontal lobe size\\[6\\]"):
        f.convolve(d_in, d_out)
    f.initialize(np.ones(3) / 5, np.ones(7) / 13)
    with pytest.raises(tucuda_amd.UnsupportedLobe, match="vertical lobe size\\[7"):
        f.convolve(d_in, d_out)
    # An output smaller than the region written is refused before any launch.
    f.initialize(np.ones(3) / 5, np.ones(3) / 5)
    small = to_dev(np.zeros((60, 64), np.float32), dev)
    with pytest.raises(tucuda_amd.TucuError):
        f.convolve(d_in, small, True)


def test_fir_strided_views(dev, kernel_family):
    """Row ranges with a stride larger than the row (slices of a bigger
    buffer), as cu::Array2 rows padded to 256 bytes are."""
    import torch
    rng = np.random.default_rng(5)
    lobeH, lobeV = random_lobe(rng, 9), random_lobe(rng, 8)
    big = synth.uniform_image(3, 90, 400)
    img = big[5:85, 7:307]
    ref = np.full((80, 300), sentinel(np.float32), np.float32)
    oracle.fir2(lobeH, lobeV, np.ascontiguousarray(img), ref, True)
    d_big = to_dev(big, dev)
    d_outbig = torch.full((100, 512), sentinel(np.float32), device=dev)
    import tucuda_amd
    tucuda_amd.FIRFilter2().initialize(lobeH, lobeV).convolve(d_big[5:85, 7:307], d_outbig[10:90, 3:303])
    got = to_host(d_outbig)
    assert_f32_parity(got[10:90, 3:303], ref, "strided")
    outside = got.copy()
    outside[10:90, 3:303] = sentinel(np.float32)
    assert (outside == sentinel(np.float32)).all()


@pytest.mark.parametrize("nH,nV", [(1, 1), (3, 5), (17, 9), (9, 17), (33, 3), (5, 33), (11, 25)])
def test_general_taps(dev, kernel_family, nH, nV):
    """General-taps extension against the oracle, every radius class and the
    edges of the tap-count range; u8 and float inputs, shift both ways."""
    import tucuda_amd
    rng = np.random.default_rng(nH * 40 + nV)
    th = rng.standard_normal(nH)
    tv = rng.standard_normal(nV)
    th, tv = (th / np.abs(th).sum()).astype(np.float32), (tv / np.abs(tv).sum()).astype(np.float32)
    for ik, shift in itertools.product(("u8", "f32"), (True, False)):
        img = _input(ik, nH + nV, 77, 301)
        ref = np.full(img.shape, sentinel(np.float32), np.float32)
        oracle.fir2_taps(th, tv, img, ref, shift)
        d_out = to_dev(np.full(img.shape, sentinel(np.float32), np.float32), dev)
        tucuda_amd.SeparableFilter2(th, tv).convolve(to_dev(img, dev), d_out, shift)
        assert_f32_parity(to_host(d_out), ref, f"taps {nH}x{nV} {ik} shift={shift}")
