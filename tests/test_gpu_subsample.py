"""cu::subsample on the GPU (TU/cu/algorithm.h:183-242, operators of
TU/cu/functional.h:58-256) against the oracle.

The kernel evaluates each operator's source expression in FP32 without
contraction, as the oracle does, and converts with the same rules, so the
bar here is bit-exact equality for every operator and element-type pair.
Cases: odd and even extents, several workgroup columns (wave-edge lanes),
unaligned input and output views (the element-load and element-store
paths), the Array2 row layout, 1080p, pixels outside the nrow/2 x ncol/2
output left untouched, the too-small no-op and the argument errors."""
import numpy as np
import pytest

import oracle
from tests._util import sentinel, to_dev, to_host
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

TYPES = {"u8": np.uint8, "f32": np.float32, "f16": np.float16, "bf16": np.uint16}
OPS = oracle.SUBSAMPLE_OPS


def _input(kind, nrow, ncol, seed=31):
    img = synth.edge_image(seed, nrow, ncol)
    if kind == "u8":
        return np.clip(img, 0, 255).astype(np.uint8)
    if kind == "f16":
        return img.astype(np.float16)
    if kind == "bf16":
        return (img.view(np.uint32) >> 16).astype(np.uint16)
    return img


def _expected(img, op, ok, out_shape):
    ref = np.full(out_shape, sentinel(TYPES[ok]), dtype=TYPES[ok])
    return oracle.subsample(img, ref, op)


def _run(dev, img, op, ok, out_shape, in_off=0, out_off=0):
    """GPU subsample of img (placed at column in_off of a wider buffer) into a
    sentinel-filled output of out_shape (at column out_off); returns the
    output as a host array."""
    import torch
    import tucuda_amd
    nrow, ncol = img.shape
    big = np.zeros((nrow, ncol + in_off), dtype=img.dtype)
    big[:, in_off:] = img
    x = to_dev(big, dev)[:, in_off:]
    fill = np.full((out_shape[0], out_shape[1] + out_off), sentinel(TYPES[ok]), dtype=TYPES[ok])
    o = to_dev(fill, dev)[:, out_off:]
    tucuda_amd.subsample(x, o, op)
    torch.cuda.synchronize()
    return to_host(o)


def _same(got, want, what):
    g = np.asarray(got)
    w = np.asarray(want)
    assert g.shape == w.shape
    if g.dtype.kind == "f":
        bad = ~((g == w) | (np.isnan(g) & np.isnan(w)))
    else:
        bad = g != w
    assert not bad.any(), (what, int(bad.sum()), np.argwhere(bad)[:3].tolist(),
                           [g[tuple(i)] for i in np.argwhere(bad)[:3]], [w[tuple(i)] for i in np.argwhere(bad)[:3]])


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("ik", ["u8", "f32", "f16", "bf16"])
@pytest.mark.parametrize("ok", ["u8", "f32", "f16", "bf16"])
@pytest.mark.parametrize("ncol", [61, 124])        # rows of 124 pixels take the 4-pixel loads, 61 the element loads
def test_type_pairs(dev, op, ik, ok, ncol):
    if ik == "u8" and op != "identity1x1":
        pytest.skip("operators other than identity1x1 need a floating-point input")
    img = _input(ik, 97, ncol)
    shape = (97 // 2 + 2, ncol // 2 + 3)           # larger than written: the rest must stay untouched
    _same(_run(dev, img, op, ok, shape), _expected(img, op, ok, shape), f"{op} {ik}->{ok}")


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("shape", [(37, 517), (64, 48), (2, 2), (3, 5), (130, 1031)])
def test_extents(dev, op, shape):
    img = _input("f32", *shape, seed=32)
    oshape = (shape[0] // 2, shape[1] // 2)
    _same(_run(dev, img, op, "f32", oshape), _expected(img, op, "f32", oshape), f"{op} {shape}")


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("offs", [(1, 0), (0, 1), (3, 1), (2, 2)])
def test_unaligned_views(dev, op, offs):
    img = _input("f32", 45, 300, seed=33)
    oshape = (22, 150)
    got = _run(dev, img, op, "f32", oshape, in_off=offs[0], out_off=offs[1])
    _same(got, _expected(img, op, "f32", oshape), f"{op} offsets {offs}")


@pytest.mark.parametrize("op", ["identity1x1", "gaussian5x5", "sobelAbs3x3"])
@pytest.mark.parametrize("ik", ["u8", "f32", "bf16"])
def test_1080p(dev, op, ik):
    if ik == "u8" and op != "identity1x1":
        pytest.skip("floating-point input only")
    img = _input(ik, 1080, 1920, seed=34)
    oshape = (540, 960)
    _same(_run(dev, img, op, "f32", oshape), _expected(img, op, "f32", oshape), f"{op} 1080p {ik}")


@pytest.mark.parametrize("op", ["identity1x1", "det3x3"])
def test_array2_layout(dev, op):
    import torch
    import tucuda_amd
    img = _input("f32", 301, 403, seed=35)
    x = tucuda_amd.array2(301, 403, torch.float32, dev)
    x.copy_(torch.from_numpy(img).to(dev))
    o = tucuda_amd.array2(150, 201, torch.float32, dev)
    tucuda_amd.subsample(x, o, op)
    torch.cuda.synchronize()
    ref = np.zeros((150, 201), np.float32)
    _same(to_host(o), oracle.subsample(img, ref, op), f"{op} Array2")


@pytest.mark.parametrize("shape", [(1, 40), (40, 1), (1, 1)])
def test_too_small_is_noop(dev, shape):
    img = _input("f32", *shape)
    got = _run(dev, img, "gaussian5x5", "f32", (4, 4))
    assert (got == sentinel(np.float32)).all()


def test_errors(dev):
    import torch
    import tucuda_amd
    x8 = torch.zeros((16, 16), dtype=torch.uint8, device=dev)
    xf = torch.zeros((16, 16), dtype=torch.float32, device=dev)
    o = torch.zeros((8, 8), dtype=torch.float32, device=dev)
    with pytest.raises(tucuda_amd.TucuError, match="floating-point"):
        tucuda_amd.subsample(x8, o, "diffH1x3")
    with pytest.raises(tucuda_amd.TucuError, match="smaller"):
        tucuda_amd.subsample(xf, o[:7], "identity1x1")
    with pytest.raises(ValueError):
        tucuda_amd.subsample(xf, o, "no_such_operator")
    with pytest.raises(tucuda_amd.TucuError, match="unknown"):
        tucuda_amd.subsample(xf, o, 12)
    tucuda_amd.subsample(x8, torch.zeros((8, 8), dtype=torch.uint8, device=dev), "identity1x1")
    torch.cuda.synchronize()
