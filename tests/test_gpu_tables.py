"""Prebuilt operand tables: the host restatement of the matrix-core kernels'
table build (mfma_host_tabs) against the build each workgroup does
(TUCU_PREBUILT_TABLES=0): every launch that takes the tables from device
memory writes the same bits as one that builds them, for the six-output
Gaussian (both radius classes, several input / output types, both shifts),
FIRFilter2 lobes and general taps (one-output mode), and ICIA's moments."""
import os

import numpy as np
import pytest

from tests._util import to_dev
from tucuda_amd import synth

pytestmark = pytest.mark.gpu


def _both(fn):
    """fn() with prebuilt tables, then with tables built in every workgroup."""
    import torch
    os.environ.pop("TUCU_PREBUILT_TABLES", None)
    a = fn()
    torch.cuda.synchronize()
    os.environ["TUCU_PREBUILT_TABLES"] = "0"
    try:
        b = fn()
        torch.cuda.synchronize()
    finally:
        os.environ.pop("TUCU_PREBUILT_TABLES", None)
    return a, b


def _bits(t):
    import torch
    return t.contiguous().view(torch.int16 if t.element_size() == 2 else
                               torch.uint8 if t.element_size() == 1 else torch.int32).cpu().numpy()


@pytest.mark.parametrize("sigma", [0.9, 2.0, 3.3])
@pytest.mark.parametrize("dt", ["float32", "bfloat16", "uint8"])
@pytest.mark.parametrize("shift", [True, False])
def test_gauss6_prebuilt_tables_bitwise(dev, sigma, dt, shift):
    import torch
    import tucuda_amd
    img = synth.edge_image(11, 300, 517)
    tdt = getattr(torch, dt)
    x = to_dev(np.clip(img, 0, 255).astype(np.uint8) if dt == "uint8" else img, dev).to(tdt)
    odt = torch.float32 if dt == "uint8" else tdt
    g = tucuda_amd.FIRGaussianConvolver2(sigma)

    def run():
        outs = [torch.full(x.shape, 7.0, dtype=odt, device=dev) for _ in range(6)]
        g.convolve6(x, outs, shift=shift)
        return outs
    a, b = _both(run)
    for k in range(6):
        assert np.array_equal(_bits(a[k]), _bits(b[k])), f"output {k}"


@pytest.mark.parametrize("nH,nV", [(9, 5), (17, 17), (4, 8)])
def test_fir_prebuilt_tables_bitwise(dev, nH, nV):
    import torch
    import tucuda_amd
    rng = np.random.default_rng(nH * 100 + nV)
    lH = np.abs(rng.standard_normal(nH)).astype(np.float32)
    lV = np.abs(rng.standard_normal(nV)).astype(np.float32)
    lH /= lH.sum()
    lV /= lV.sum()
    x = to_dev(synth.u8_image(5, 640, 700), dev)
    tH = rng.standard_normal(17).astype(np.float32)
    tV = rng.standard_normal(11).astype(np.float32)

    def run():
        o1 = torch.zeros(x.shape, device=dev)
        tucuda_amd.FIRFilter2().initialize(lH, lV).convolve(x, o1)
        o2 = torch.zeros(x.shape, dtype=torch.float16, device=dev)
        tucuda_amd.SeparableFilter2(tH / np.abs(tH).sum(), tV / np.abs(tV).sum()).convolve(x, o2)
        return o1, o2
    a, b = _both(run)
    assert np.array_equal(_bits(a[0]), _bits(b[0]))
    assert np.array_equal(_bits(a[1]), _bits(b[1]))


@pytest.mark.parametrize("model", ["affine", "projective"])
@pytest.mark.parametrize("sigma", [1.5, 3.0])
def test_icia_prebuilt_tables_bitwise(dev, model, sigma):
    import tucuda_amd
    x = to_dev(synth.edge_image(12, 257, 391), dev)
    a, b = _both(lambda: tucuda_amd.icia_moment(sigma, x, model))
    assert np.array_equal(_bits(a), _bits(b))


def test_more_lobe_sets_than_table_slots(dev):
    """Past the 64 table slots of a device, launches build their tables in
    every workgroup: 70 distinct sigmas, each result equal to the in-kernel
    build's bits."""
    import torch
    import tucuda_amd
    x = to_dev(synth.edge_image(13, 96, 200), dev)
    for i in range(70):
        sigma = 0.8 + 0.047 * i
        g = tucuda_amd.FIRGaussianConvolver2(sigma)

        def run():
            o = torch.zeros_like(x)
            g.smooth(x, o)
            return o
        a, b = _both(run)
        assert np.array_equal(_bits(a), _bits(b)), sigma


def test_graph_captured_before_the_lobes_were_seen(dev):
    """A launch captured into a HIP graph before its lobes were ever used
    gets no prebuilt table (nothing is uploaded under capture) and builds
    its tables itself: every replay equals the eager launch's bits, with
    eager work in between."""
    import torch
    import tucuda_amd
    x = to_dev(synth.edge_image(14, 130, 260), dev)
    g = tucuda_amd.FIRGaussianConvolver2(2.71828)      # a sigma no other test uses
    out = torch.zeros_like(x)
    st = torch.cuda.Stream(dev)
    st.wait_stream(torch.cuda.current_stream(dev))
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            g.convolve6(x, {"smooth": out})
    torch.cuda.current_stream(dev).wait_stream(st)
    ref = torch.zeros_like(x)
    g.convolve6(x, {"smooth": ref})                  # eager: uploads the table now
    torch.cuda.synchronize()
    for _ in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(_bits(out), _bits(ref))
        tmp = torch.zeros_like(x)
        g.convolve6(x, {"diffH": tmp})               # eager work between replays
        torch.cuda.synchronize()
