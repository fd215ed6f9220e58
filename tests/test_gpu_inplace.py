"""In-place calls.  The reference's FIRFilter2::convolve filters through its
intermediate _buf (FIRFilter.h:298-377), so an output that is the input
itself is well defined there; the fused launches read the input's halo
while other workgroups write, so the C ABI filters such a call from a copy
of the input.  Each in-place result must equal the out-of-place one bit for
bit; outputs of one fused call that overlap each other are refused."""
import numpy as np
import pytest

from tests._util import to_dev, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shift", [True, False])
def test_fir2_in_place(dev, shift):
    import torch
    import tucuda_amd
    rng = np.random.default_rng(5)
    lobeH = rng.standard_normal(9).astype(np.float32)
    lobeV = rng.standard_normal(16).astype(np.float32)
    img = synth.edge_image(3, 130, 300)
    f = tucuda_amd.FIRFilter2().initialize(lobeH, lobeV)
    x = to_dev(img, dev)
    want = x.clone()                                  # out-of-place over a copy of the input
    f.convolve(x.clone(), want, shift)
    f.convolve(x, x, shift)                           # in place
    torch.cuda.synchronize()
    assert torch.equal(x, want)


def test_general_taps_in_place(dev):
    import torch
    import tucuda_amd
    rng = np.random.default_rng(6)
    th, tv = rng.standard_normal(13).astype(np.float32), rng.standard_normal(7).astype(np.float32)
    x = to_dev(synth.edge_image(4, 90, 210), dev)
    f = tucuda_amd.SeparableFilter2(th, tv)
    want = x.clone()
    f.convolve(x.clone(), want)
    f.convolve(x, x)
    torch.cuda.synchronize()
    assert torch.equal(x, want)


@pytest.mark.parametrize("method", ["smooth", "diffHV"])
def test_gauss_method_in_place(dev, method):
    import torch
    import tucuda_amd
    g = tucuda_amd.FIRGaussianConvolver2(2.5)
    x = to_dev(synth.edge_image(5, 111, 257), dev)
    want = x.clone()
    getattr(g, method)(x.clone(), want)
    getattr(g, method)(x, x)
    torch.cuda.synchronize()
    assert torch.equal(x, want)


def test_convolve6_one_output_is_the_input(dev):
    import torch
    import tucuda_amd
    g = tucuda_amd.FIRGaussianConvolver2(3.0)
    x = to_dev(synth.edge_image(6, 120, 240), dev)
    outs = {m: torch.zeros_like(x) for m in tucuda_amd.METHOD_ORDER}
    want = {m: torch.zeros_like(x) for m in tucuda_amd.METHOD_ORDER}
    want["diffV"] = x.clone()              # the border keeps the input, as in the in-place call
    g.convolve6(x.clone(), want)
    outs["diffV"] = x
    g.convolve6(x, outs)
    torch.cuda.synchronize()
    for m in tucuda_amd.METHOD_ORDER:
        assert torch.equal(outs[m], want[m]), m


def test_overlapping_outputs_refused(dev):
    import torch
    import tucuda_amd
    g = tucuda_amd.FIRGaussianConvolver2(2.0)
    x = to_dev(synth.edge_image(7, 64, 96), dev)
    o = torch.zeros_like(x)
    with pytest.raises(tucuda_amd.TucuError):
        g.convolve6(x, {"smooth": o, "diffH": o})


def test_subsample_and_pyramid_in_place(dev):
    import torch
    import tucuda_amd
    img = synth.edge_image(8, 100, 150)
    x = to_dev(img, dev)
    want = torch.zeros_like(x)
    tucuda_amd.subsample(x.clone(), want, "gaussian5x5")
    y = x.clone()
    tucuda_amd.subsample(y, y, "gaussian5x5")
    torch.cuda.synchronize()
    assert torch.equal(y[:50, :75], want[:50, :75])
    g = tucuda_amd.FIRGaussianConvolver2(1.5)
    want = torch.zeros_like(x)
    g.smoothSubsample(x.clone(), want)
    y = x.clone()
    g.smoothSubsample(y, y)
    torch.cuda.synchronize()
    assert torch.equal(y[:50, :75], want[:50, :75])


def test_batch_refuses_in_place_jobs(dev):
    import torch
    import tucuda_amd
    x = to_dev(synth.edge_image(9, 64, 64), dev)
    with pytest.raises(tucuda_amd.TucuError):
        tucuda_amd.GaussBatch([(x, {"smooth": x}, 2.0)])
