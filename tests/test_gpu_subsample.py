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


# operators other than identity1x1 need a floating-point input (a u_char
# input is refused, tested in test_errors)
OP_INPUTS = [(op, ik) for op in OPS for ik in ("u8", "f32", "f16", "bf16")
             if ik != "u8" or op == "identity1x1"]


@pytest.mark.parametrize("op,ik", OP_INPUTS)
@pytest.mark.parametrize("ok", ["u8", "f32", "f16", "bf16"])
@pytest.mark.parametrize("ncol", [61, 124])        # rows of 124 pixels take the 4-pixel loads, 61 the element loads
def test_type_pairs(dev, op, ik, ok, ncol):
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


@pytest.mark.parametrize("op,ik", [(op, ik) for op in ("identity1x1", "gaussian5x5", "sobelAbs3x3")
                                   for ik in ("u8", "f32", "bf16") if ik != "u8" or op == "identity1x1"])
def test_1080p(dev, op, ik):
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


# ---------------------------------------------------------------------------
# The fused pyramid step (tucu_gauss2_pyramid, FIRGaussianConvolver2::
# smoothSubsample): equal to smooth into a zero-filled image of the output's
# element type followed by subsample(identity1x1), to the FP32 bar.

def _pyr_expected(img, sigma, rule, ok, shift):
    import tucuda_amd
    if rule == 0:
        lobe = oracle.gauss_lobes(sigma)[0]
        smooth = lambda out: oracle.fir2(lobe, lobe, img, out, shift)          # noqa: E731
    else:   # ceil(4 sigma) radii: the same symmetric kernel (lobe 0 = r + 1 entries, centre last) as general taps
        lobe = tucuda_amd.FIRGaussianConvolver2.lobes(sigma, rule)[0]
        taps = np.concatenate([lobe[:-1], lobe[-1:], lobe[:-1][::-1]]).astype(np.float32)
        smooth = lambda out: oracle.fir2_taps(taps, taps, img, out, shift)     # noqa: E731
    nrow, ncol = img.shape
    full_f = np.zeros((nrow, ncol), np.float32)
    smooth(full_f)
    half_f = np.zeros((nrow // 2, ncol // 2), np.float32)
    oracle.subsample(full_f, half_f, "identity1x1")
    full = np.zeros((nrow, ncol), TYPES[ok])
    smooth(full)
    half = np.full((nrow // 2, ncol // 2), sentinel(TYPES[ok]), TYPES[ok])
    oracle.subsample(full, half, "identity1x1")
    return half, half_f


def _pyr_check(got, ref, ref_f, ok, what):
    from tests._util import assert_f32_parity, assert_u8_parity
    if ok == "f32":
        assert_f32_parity(got, ref, what)
    elif ok == "u8":
        assert_u8_parity(got, ref, ref_f, what)
    else:
        g = (got.astype(np.uint32) << 16).view(np.float32) if ok == "bf16" else got.astype(np.float32)
        r = (ref.astype(np.uint32) << 16).view(np.float32) if ok == "bf16" else ref.astype(np.float32)
        ulp = np.abs(r.astype(np.float64)) * (2.0 ** -10 if ok == "f16" else 2.0 ** -7)
        tol = np.maximum(1e-5 * np.abs(ref_f), 5e-6 * 255) + ulp
        bad = ~(np.abs(g.astype(np.float64) - r) <= tol)
        assert not bad.any(), (what, int(bad.sum()), np.argwhere(bad)[:3].tolist())


def _pyr_run(dev, img, sigma, rule, ok, shift, in_off=0):
    import torch
    import tucuda_amd
    nrow, ncol = img.shape
    big = np.zeros((nrow, ncol + in_off), dtype=img.dtype)
    big[:, in_off:] = img
    x = to_dev(big, dev)[:, in_off:]
    o = to_dev(np.full((nrow // 2 + 1, ncol // 2 + 2), sentinel(TYPES[ok]), TYPES[ok]), dev)
    g = tucuda_amd.FIRGaussianConvolver2(sigma, rule)
    g.smoothSubsample(x, o, shift)
    torch.cuda.synchronize()
    got = to_host(o)
    fill = np.full(1, sentinel(TYPES[ok]), TYPES[ok])[0]      # the sentinel as stored in the output type
    assert (got[nrow // 2:, :] == fill).all() and (got[:, ncol // 2:] == fill).all()
    return got[:nrow // 2, :ncol // 2]


@pytest.mark.parametrize("sigma,rule", [(0.8, 0), (1.0, 0), (2.0, 0), (3.0, 0), (0.8, 1), (1.3, 1), (3.1, 1)])
@pytest.mark.parametrize("shift", [True, False])
@pytest.mark.parametrize("shape", [(97, 61), (300, 250), (37, 517), (9, 40)])
def test_pyramid_step(dev, sigma, rule, shift, shape):
    img = _input("f32", *shape, seed=36)
    ref, ref_f = _pyr_expected(img, sigma, rule, "f32", shift)
    _pyr_check(_pyr_run(dev, img, sigma, rule, "f32", shift), ref, ref_f, "f32",
               f"pyramid sigma={sigma} rule={rule} shift={shift} {shape}")


@pytest.mark.parametrize("ik", ["u8", "f32", "f16", "bf16"])
@pytest.mark.parametrize("ok", ["u8", "f32", "f16", "bf16"])
def test_pyramid_type_pairs(dev, ik, ok):
    img = _input(ik, 130, 250, seed=37)
    ref, ref_f = _pyr_expected(img, 2.0, 0, ok, True)
    _pyr_check(_pyr_run(dev, img, 2.0, 0, ok, True, in_off=1 if ik != "u8" else 3), ref, ref_f, ok,
               f"pyramid {ik}->{ok}")


@pytest.mark.parametrize("sigma", [1.0, 3.0])
def test_pyramid_1080p_equals_two_step(dev, sigma):
    """At full size: the fused step against smooth + subsample on the GPU
    (same lobes, both paths FP32) and against the oracle."""
    import torch
    import tucuda_amd
    img = _input("f32", 1080, 1920, seed=38)
    x = to_dev(img, dev)
    g = tucuda_amd.FIRGaussianConvolver2(sigma)
    full = torch.zeros_like(x)
    g.smooth(x, full)
    two = torch.zeros((540, 960), dtype=torch.float32, device=dev)
    tucuda_amd.subsample(full, two, "identity1x1")
    one = torch.full((540, 960), float("nan"), dtype=torch.float32, device=dev)
    g.smoothSubsample(x, one)
    torch.cuda.synchronize()
    ref, ref_f = _pyr_expected(img, sigma, 0, "f32", True)
    _pyr_check(to_host(one), ref, ref_f, "f32", f"pyramid 1080p sigma={sigma}")
    _pyr_check(to_host(one), to_host(two), to_host(two), "f32", f"pyramid fused vs two-step sigma={sigma}")


def test_graph_replays_equal_eager(dev):
    """subsample and the fused pyramid step captured in a HIP graph: every
    replay (with eager work and synchronisations between them, the inputs
    changed between replays) equals the eager launches bit for bit."""
    import torch
    import tucuda_amd
    g = tucuda_amd.FIRGaussianConvolver2(2.0)
    x = to_dev(_input("f32", 301, 403, seed=39), dev)
    half_a = torch.zeros((150, 201), dtype=torch.float32, device=dev)
    half_b = torch.zeros_like(half_a)

    def work():
        tucuda_amd.subsample(x, half_a, "gaussian5x5")
        g.smoothSubsample(x, half_b)

    work()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    st = torch.cuda.Stream(dev)
    st.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(st):
        with torch.cuda.graph(graph, stream=st):
            work()
    torch.cuda.current_stream(dev).wait_stream(st)
    for rep in range(4):
        x.copy_(to_dev(_input("f32", 301, 403, seed=40 + rep), dev))
        half_a.fill_(float("nan"))
        half_b.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got_a, got_b = to_host(half_a), to_host(half_b)
        work()                                                # eager, same inputs
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got_a, to_host(half_a), err_msg=f"subsample replay {rep}")
        np.testing.assert_array_equal(got_b, to_host(half_b), err_msg=f"pyramid replay {rep}")
