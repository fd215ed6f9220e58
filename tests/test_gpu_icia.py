"""ICIA's error moment (ICIA.h:346-369): the fused diffH / diffV pass that
feeds the moment reduction without writing the gradients, against the
oracle -- the FP32-faithful gradients of FIRGaussianConvolver2 in
zero-filled images, summed in FP64 (oracle.icia_moment, pinned by closed
forms in tests/test_oracle_icia.py).  Tolerance: each moment within 5e-5 of
the sum of its terms' magnitudes (the gradients agree to about 1e-6
relative; a product of two doubles that; the sums' order differs)."""
import ctypes

import numpy as np
import pytest

import oracle
from tests._util import to_dev
from tucuda_amd import synth

pytestmark = pytest.mark.gpu


def _oracle_moment(img, sigma, model):
    lobes = oracle.gauss_lobes(sigma)
    eH = np.zeros(img.shape, np.float32)
    eV = np.zeros(img.shape, np.float32)
    oracle.gauss("diffH", sigma, img, eH, lobes=lobes)
    oracle.gauss("diffV", sigma, img, eV, lobes=lobes)
    return oracle.icia_moment(eH, eV, model)


@pytest.mark.parametrize("model", ["affine", "projective"])
@pytest.mark.parametrize("sigma", [1.0, 2.0, 3.0])
@pytest.mark.parametrize("shape", [(120, 200), (257, 131)])
def test_icia_moment_matches_oracle(dev, model, sigma, shape):
    import torch
    import tucuda_amd
    img = synth.edge_image(int(sigma * 10) + shape[0], *shape)
    got = tucuda_amd.icia_moment(sigma, to_dev(img, dev), model).cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    want, absm = _oracle_moment(img, sigma, model)
    assert got.shape == want.shape
    np.testing.assert_array_less(np.abs(got - want), 5e-5 * absm + 1e-30)


def test_icia_moment_full_hd_and_u8(dev):
    import torch
    import tucuda_amd
    sigma = 3.0
    img = synth.edge_image(5, 1080, 1920)
    got = tucuda_amd.icia_moment(sigma, to_dev(img, dev), "projective").cpu().numpy().astype(np.float64)
    want, absm = _oracle_moment(img, sigma, "projective")
    np.testing.assert_array_less(np.abs(got - want), 5e-5 * absm)
    img8 = synth.u8_image(6, 300, 420)
    got = tucuda_amd.icia_moment(2.0, to_dev(img8, dev), "affine").cpu().numpy().astype(np.float64)
    want, absm = _oracle_moment(img8, 2.0, "affine")
    np.testing.assert_array_less(np.abs(got - want), 5e-5 * absm)
    torch.cuda.synchronize()


def test_icia_moment_of_the_fused_gradients(dev):
    """The moment kernel sums exactly the gradients the fused six-output
    launch writes: the oracle's moment of those images agrees to 1e-5."""
    import torch
    import tucuda_amd
    sigma = 2.5
    img = synth.edge_image(8, 333, 517)
    x = to_dev(img, dev)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    eH, eV = torch.zeros_like(x), torch.zeros_like(x)
    g.convolve6(x, {"diffH": eH, "diffV": eV})
    got = tucuda_amd.icia_moment(sigma, x, "affine").cpu().numpy().astype(np.float64)
    want, absm = oracle.icia_moment(eH.cpu().numpy(), eV.cpu().numpy(), "affine")
    np.testing.assert_array_less(np.abs(got - want), 1e-5 * absm)


def test_icia_moment_small_image_and_workspace(dev):
    """An image too small for the lobes leaves both gradients zero (the
    reference's convolve returns early): the moment is exactly 0.  The
    workspace query and a too-small workspace."""
    import torch
    import tucuda_amd
    from tucuda_amd import _capi
    from tucuda_amd.filters import image_desc
    x = torch.ones((20, 20), device=dev)
    got = tucuda_amd.icia_moment(3.0, x, "affine")
    assert (got.cpu().numpy() == 0).all()
    l0, l1, _ = tucuda_amd.gauss_lobes(2.0)
    fp = ctypes.POINTER(ctypes.c_float)
    y = torch.ones((200, 300), device=dev)
    src = image_desc(y)
    need = ctypes.c_size_t(0)
    lib = _capi.load()
    assert lib.tucu_icia_moment(l0.ctypes.data_as(fp), l0.size, l1.ctypes.data_as(fp), l1.size, ctypes.byref(src),
                                6, None, None, ctypes.byref(need), None) == 0
    assert need.value > 0
    small = ctypes.c_size_t(need.value - 256)
    ws = torch.empty(need.value, dtype=torch.uint8, device=dev)
    mom = torch.empty(21, device=dev)
    rc = lib.tucu_icia_moment(l0.ctypes.data_as(fp), l0.size, l1.ctypes.data_as(fp), l1.size, ctypes.byref(src),
                              6, ctypes.c_void_p(mom.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                              ctypes.byref(small), None)
    assert rc == _capi.TUCU_ERR_INVALID
    # a constant image has zero gradients everywhere: the moment is 0
    got = tucuda_amd.icia_moment(2.0, y, "projective")
    assert float(got.abs().max()) == 0.0


def test_icia_moment_many_partials_deterministic(dev):
    """A 4096 x 8192 u8 image: more workgroup partials than the reduction's
    lanes take in one pass of loads (its loop runs more than once).  The
    moment agrees with the oracle's moment of the gradients the fused launch
    writes, and two calls give the same bits (a fixed summation order)."""
    import torch
    import tucuda_amd
    sigma = 2.0
    img = synth.u8_image(9, 4096, 8192)
    x = to_dev(img, dev)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    eH = torch.zeros(x.shape, dtype=torch.float32, device=dev)
    eV = torch.zeros_like(eH)
    g.convolve6(x, {"diffH": eH, "diffV": eV})
    a = tucuda_amd.icia_moment(sigma, x, "projective")
    b = tucuda_amd.icia_moment(sigma, x, "projective")
    got = a.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), b.cpu().numpy().view(np.uint32))
    want, absm = oracle.icia_moment(eH.cpu().numpy(), eV.cpu().numpy(), "projective")
    np.testing.assert_array_less(np.abs(got.astype(np.float64) - want), 1e-5 * absm)


def test_icia_moment_graph_replay(dev):
    """tucu_icia_moment captured in a HIP graph (the moment pass and the
    reduction; the workspace and the moment come from the graph's pool):
    every replay gives the bits of the eager call on the image the input
    buffer holds at that replay, with eager calls between the replays."""
    import torch
    import tucuda_amd
    sigma = 3.0
    imgs = [synth.edge_image(20 + i, 1080, 1920) for i in range(3)]
    x = to_dev(imgs[0], dev)
    refs = []
    for img in imgs:                                  # eager; uploads the operand tables
        refs.append(tucuda_amd.icia_moment(sigma, to_dev(img, dev), "projective").cpu().numpy())
    st = torch.cuda.Stream(dev)
    st.wait_stream(torch.cuda.current_stream(dev))
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            m = tucuda_amd.icia_moment(sigma, x, "projective")
    torch.cuda.current_stream(dev).wait_stream(st)
    torch.cuda.synchronize()
    for rep in range(6):
        i = rep % 3
        x.copy_(to_dev(imgs[i], dev))
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(m.cpu().numpy().view(np.uint32), refs[i].view(np.uint32)), rep
        tucuda_amd.icia_moment(2.0, to_dev(imgs[(i + 1) % 3], dev), "affine")   # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("model", ["affine", "projective"])
@pytest.mark.parametrize("sigma", [2.0, 3.0])
@pytest.mark.parametrize("ik", ["u8", "f16", "bf16"])
def test_icia_moment_narrow_inputs(dev, ik, sigma, model):
    """u8, float16 and bfloat16 images, both radius classes and both maps,
    against the oracle's moment of the same values as float32 (16-bit
    values are exact in float32)."""
    import tucuda_amd
    img = synth.edge_image(30, 257, 331)
    if ik == "u8":
        img = dev_img = np.clip(img, 0, 255).astype(np.uint8)
    elif ik == "f16":
        dev_img = img.astype(np.float16)
        img = dev_img.astype(np.float32)
    else:
        dev_img = (img.view(np.uint32) >> 16).astype(np.uint16)      # truncation: exact bf16 values
        img = (dev_img.astype(np.uint32) << 16).view(np.float32)
    got = tucuda_amd.icia_moment(sigma, to_dev(dev_img, dev), model).cpu().numpy().astype(np.float64)
    want, absm = _oracle_moment(img, sigma, model)
    np.testing.assert_array_less(np.abs(got - want), 5e-5 * absm + 1e-30)
