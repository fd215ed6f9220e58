"""The GPU kernels against the committed golden fixtures (tests/golden):
the six Gaussian methods, fused and one by one, on both kernel families;
FIRFilter2 with four lobe-length pairs and both shifts; cu::subsample with
every operator (bit-exact)."""
import os

import numpy as np
import pytest

from tests._util import assert_f32_parity, to_dev, to_host

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
METHODS = ("smooth", "diffH", "diffV", "diffHH", "diffHV", "diffVV")


def _npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.mark.parametrize("name", ["u8_64x48", "f32_97x61"])
@pytest.mark.parametrize("sigma", [0.8, 1.2, 2.0, 2.5, 3.0])
def test_gauss(dev, kernel_family, name, sigma):
    import torch
    import tucuda_amd
    g = _npz("gauss.npz")
    x = to_dev(g[name], dev)
    conv = tucuda_amd.FIRGaussianConvolver2(sigma)
    fused = {m: torch.zeros(x.shape, dtype=torch.float32, device=dev) for m in METHODS}
    conv.convolve6(x, fused)
    single = {m: torch.zeros(x.shape, dtype=torch.float32, device=dev) for m in METHODS}
    for m in METHODS:
        getattr(conv, m)(x, single[m])
    torch.cuda.synchronize()
    for m in METHODS:
        want = g[f"{name}_s{sigma}_{m}"]
        assert_f32_parity(to_host(fused[m]), want, f"fused {name} {sigma} {m}")
        assert_f32_parity(to_host(single[m]), want, f"{m} {name} {sigma}")


@pytest.mark.parametrize("k", range(4))
@pytest.mark.parametrize("shift", [True, False])
def test_fir(dev, kernel_family, k, shift):
    import torch
    import tucuda_amd
    f = _npz("fir.npz")
    x = to_dev(f["image"], dev)
    out = torch.zeros(x.shape, dtype=torch.float32, device=dev)
    tucuda_amd.FIRFilter2().initialize(f[f"p{k}_lobeH"], f[f"p{k}_lobeV"]).convolve(x, out, shift)
    torch.cuda.synchronize()
    assert_f32_parity(to_host(out), f[f"p{k}_shift{int(shift)}"], f"fir pair {k} shift {shift}")


def test_subsample(dev):
    import torch
    import tucuda_amd
    s = _npz("subsample.npz")
    x = to_dev(s["image"], dev)
    for op in tucuda_amd.SUBSAMPLE_OPS:
        o = torch.zeros((18, 25), dtype=torch.float32, device=dev)
        tucuda_amd.subsample(x, o, op)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(to_host(o), s[op], err_msg=op)
    o = torch.zeros((18, 25), dtype=torch.uint8, device=dev)
    tucuda_amd.subsample(to_dev(s["image_u8"], dev), o, "identity1x1")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(to_host(o), s["identity1x1_u8"])
