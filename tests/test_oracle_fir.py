"""Pins the oracle's separable FIR (FIRFilter.h:37-40, 66-231, 286-287,
298-377) against independent statements of what the filter must compute:
a brute-force 2-D correlation in float64, impulse responses (orientation,
tap pairing and output offset), exact responses to constant / linear /
quadratic images, the reference's own CPU FIR convention
(Testprog/filterTest/filterImageGold.h:18-33), the untouched border and the
size / error rules."""
import itertools

import numpy as np
import pytest

import oracle
from tests._util import random_lobe, sentinel
from tucuda_amd import synth

SIZES = oracle.SUPPORTED_LOBES


def brute_force(lobeH, lobeV, img, shift=True, fill=np.nan):
    """out[y+oy][x+ox] = sum_{i,j} kv[i] kh[j] img[y+i][x+j] over the valid
    region, in float64, straight from the 2-D window definition."""
    kh, kv = oracle.full_kernel(lobeH), oracle.full_kernel(lobeV)
    rH, rV = (kh.size - 1) // 2, (kv.size - 1) // 2
    nrow, ncol = img.shape
    out = np.full((nrow, ncol), fill)
    if nrow <= 2 * rV or ncol <= 2 * rH:
        return out
    f = img.astype(np.float64)
    acc = np.zeros((nrow - 2 * rV, ncol - 2 * rH))
    for i in range(kv.size):
        for j in range(kh.size):
            acc += kv[i] * kh[j] * f[i:i + nrow - 2 * rV, j:j + ncol - 2 * rH]
    oy, ox = (rV, rH) if shift else (0, 0)
    out[oy:oy + acc.shape[0], ox:ox + acc.shape[1]] = acc
    return out


def run(lobeH, lobeV, img, out_dtype=np.float32, shift=True, precision="f32"):
    out = np.full(img.shape, sentinel(out_dtype), dtype=out_dtype)
    oracle.fir2(lobeH, lobeV, img, out, shift, precision)
    return out


@pytest.mark.parametrize("nH,nV", list(itertools.product(SIZES, SIZES)))
def test_brute_force_all_pairs(nH, nV):
    rng = np.random.default_rng(nH * 100 + nV)
    lobeH, lobeV = random_lobe(rng, nH), random_lobe(rng, nV)
    img = synth.uniform_image(nH + nV, 41, 45)
    for shift in (True, False):
        ref = brute_force(lobeH, lobeV, img, shift)
        valid = ~np.isnan(ref)
        got32 = run(lobeH, lobeV, img, np.float32, shift)
        got64 = run(lobeH, lobeV, img, np.float32, shift, "f64")
        np.testing.assert_allclose(got32[valid], ref[valid], rtol=1e-5, atol=2e-4)
        np.testing.assert_allclose(got64[valid], ref[valid].astype(np.float32), rtol=1e-6, atol=1e-5)
        assert (got32[~valid] == sentinel(np.float32)).all()


@pytest.mark.parametrize("nH,nV", [(17, 17), (16, 9), (3, 2), (8, 5)])
def test_impulse_response(nH, nV):
    """A unit impulse at (y0, x0) yields kv[y0-y] * kh[x0-x] at output
    (y+oy, x+ox): the correlation of the window with the taps, tap i of a
    window weighted lobe[i] for i < r (FIRFilter.h:66-176)."""
    rng = np.random.default_rng(7)
    lobeH, lobeV = random_lobe(rng, nH), random_lobe(rng, nV)
    kh, kv = oracle.full_kernel(lobeH), oracle.full_kernel(lobeV)
    rH, rV = nH & ~1, nV & ~1
    img = np.zeros((2 * rV + 20, 2 * rH + 24), np.float32)
    y0, x0 = rV + 7, rH + 9
    img[y0, x0] = 1.0
    out = run(lobeH, lobeV, img, shift=True)
    for y in range(img.shape[0] - 2 * rV):
        for x in range(img.shape[1] - 2 * rH):
            i, j = y0 - y, x0 - x
            want = kv[i] * kh[j] if 0 <= i < kv.size and 0 <= j < kh.size else 0.0
            assert abs(out[y + rV, x + rH] - want) <= 1e-7 + 1e-6 * abs(want), (y, x)


def test_filter_image_gold_convention():
    """The reference's own CPU FIR (filterImageGold.h:18-33) writes, for a
    coefficient array coeff[0..tail] (centre last),
        val = coeff[tail]*row[j] + sum_{n<tail} coeff[n]*(row[j-tail+n] + row[j+tail-n])
    at the window centre j; the GPU filter with lobe = coeff and shift=true
    must put the same value at the same place."""
    rng = np.random.default_rng(3)
    for L in (17, 9, 5, 3):
        coeff = np.abs(rng.standard_normal(L)).astype(np.float32)
        coeff /= coeff[-1] + 2 * coeff[:-1].sum()
        tail = L - 1
        row = synth.uniform_image(L, 1, 80)[0].astype(np.float64)
        gold = np.full(80, np.nan)
        for j in range(tail, 80 - tail):
            gold[j] = coeff[tail] * row[j] + sum(coeff[n] * (row[j - tail + n] + row[j + tail - n])
                                                for n in range(tail))
        img = np.repeat(row[None, :], 7, axis=0).astype(np.float32)
        delta = np.array([0, 0, 1], np.float32)          # vertical identity (3-tap lobe, r = 2)
        out = run(coeff, delta, img)
        np.testing.assert_allclose(out[2, tail:80 - tail], gold[tail:80 - tail], rtol=1e-6)


@pytest.mark.parametrize("sigma", [1.0, 2.0, 2.5, 3.0])
def test_constant_linear_quadratic_images(sigma):
    """Exact responses of the six methods (FIRGaussianConvolver.h:178-242):
    constant c -> smooth c, first derivatives 0 (exactly), second derivatives
    c * DC(lobe2); x-ramp -> diffH = ramp gain, smooth = ramp; x^2 ->
    diffHH = 2 * sign; y-ramp -> diffV = ramp gain."""
    l0, l1, l2 = oracle.gauss_lobes(sigma, "f64")
    k1, k2 = oracle.full_kernel(l1), oracle.full_kernel(l2)
    gain1 = float(np.dot(k1, np.arange(k1.size)))
    c2 = (k2.size - 1) // 2
    quad2 = float(np.dot(k2, (np.arange(k2.size) - c2) ** 2.0))
    dc2 = float(k2.sum())
    nrow, ncol = 60, 70
    lobes32 = oracle.gauss_lobes(sigma)

    def method(m, img, precision="f64"):
        out = np.full(img.shape, np.nan, np.float32)
        lob = oracle.gauss_lobes(sigma, precision) if precision == "f64" else lobes32
        return oracle.gauss(m, sigma, img, out, True, precision, lobes=lob)

    const = np.full((nrow, ncol), 100.0, np.float32)
    for prec, tol in (("f64", 1e-4), ("f32", 5e-4)):
        for m in ("diffH", "diffV", "diffHV"):
            v = method(m, const, prec)
            assert np.nanmax(np.abs(v)) == 0.0, m          # paired differences cancel exactly
        v = method("smooth", const, prec)
        assert np.nanmax(np.abs(v - 100.0)) < tol
        for m in ("diffHH", "diffVV"):
            v = method(m, const, prec)
            assert np.nanmax(np.abs(v - 100.0 * dc2)) < tol * 10, m
    xr = np.tile(np.arange(ncol, dtype=np.float32), (nrow, 1))
    v = method("diffH", xr)
    assert np.nanmax(np.abs(v - gain1)) < 1e-4
    v = method("smooth", xr)
    valid = ~np.isnan(v)
    assert np.max(np.abs(v[valid] - xr[valid])) < 1e-3
    yr = np.tile(np.arange(nrow, dtype=np.float32)[:, None], (1, ncol))
    v = method("diffV", yr)
    assert np.nanmax(np.abs(v - gain1)) < 1e-4
    xq = (xr - ncol / 2) ** 2
    v = method("diffHH", xq.astype(np.float32))
    # second difference of x^2 is quad2 plus DC(lobe2) * x^2 (nonzero DC)
    valid = ~np.isnan(v)
    want = quad2 + dc2 * xq
    assert np.max(np.abs(v[valid] - want[valid])) < 1e-3 * np.max(np.abs(want[valid]))


def test_border_untouched_and_early_return():
    lobe = random_lobe(np.random.default_rng(0), 17)
    img = synth.uniform_image(5, 50, 60)
    out = run(lobe, lobe, img)
    assert (out[:16] == sentinel(np.float32)).all() and (out[-16:] == sentinel(np.float32)).all()
    assert (out[:, :16] == sentinel(np.float32)).all() and (out[:, -16:] == sentinel(np.float32)).all()
    assert (out[16:-16, 16:-16] != sentinel(np.float32)).all()
    for shape in [(32, 100), (100, 32), (33, 32)]:            # nrow <= 2 rV or ncol <= 2 rH
        o = run(lobe, lobe, synth.uniform_image(1, *shape))
        assert (o == sentinel(np.float32)).all()
    o = run(lobe, lobe, synth.uniform_image(1, 33, 33))
    assert o[16, 16] != sentinel(np.float32) and (o != sentinel(np.float32)).sum() == 1


def test_errors():
    img = synth.uniform_image(1, 40, 40)
    out = np.zeros_like(img)
    with pytest.raises(oracle.LobeTooLarge):
        oracle.fir2(np.ones(18), np.ones(3), img, out)
    for bad in (1, 6, 7, 10, 15):
        with pytest.raises(oracle.UnsupportedLobe):
            oracle.fir2(np.ones(bad), np.ones(3), img, out)
        with pytest.raises(oracle.UnsupportedLobe):
            oracle.fir2(np.ones(3), np.ones(bad), img, out)
    # the too-small test comes first (FIRFilter.h:304-312 precede the switch)
    oracle.fir2(np.ones(6), np.ones(3), synth.uniform_image(1, 4, 4), np.zeros((4, 4), np.float32))


def test_u8_input_and_output():
    """u_char in is converted exactly (algorithm.h:119-128); the float result
    stored into u_char saturates and truncates toward zero."""
    delta = np.array([0, 0, 1], np.float32)                      # identity lobe (r = 2)
    img = synth.u8_image(9, 20, 30)
    out = run(delta, delta, img, np.float32)
    np.testing.assert_array_equal(out[2:-2, 2:-2], img[2:-2, 2:-2].astype(np.float32))
    # Zero-weight taps still multiply their pixels (0 * NaN = NaN in the
    # reference too), so the values sit 5 columns apart.
    vals = [-5.0, 0.0, 0.4, 0.99, 1.0, 127.5, 254.999, 255.0, 255.7, 1e9, np.nan]
    img = np.zeros((5, 5 * len(vals)), np.float32)
    img[:, 2::5] = np.array(vals, np.float32)
    out = run(delta, delta, img, np.uint8)
    assert out[2, 2::5].tolist() == [0, 0, 0, 0, 1, 127, 254, 255, 255, 255, 0]


def test_conversions_match_numpy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-9, 6, 2000),
                         [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e-8, 6.1e-5, 5.96e-8, 2.98e-8,
                          np.inf, -np.inf]]).astype(np.float32)
    for x in xs:
        want = np.float32(x).astype(np.float16).view(np.uint16)
        assert oracle.float_to_half(float(x)) == int(want), x
        assert oracle.half_to_float(int(want)) == float(np.float16(x)), x
    assert np.isnan(oracle.half_to_float(oracle.float_to_half(float("nan"))))
    # bfloat16: round to nearest even on the top 16 bits
    for x in xs:
        bits = int(np.float32(x).view(np.uint32))
        lsb = (bits >> 16) & 1
        want = ((bits + 0x7FFF + lsb) >> 16) & 0xFFFF
        assert oracle.float_to_bf16(float(x)) == want, x
        assert oracle.bf16_to_float(want) == float(np.uint32(want << 16).view(np.float32))


@pytest.mark.parametrize("nH,nV", [(17, 9), (33, 33), (1, 5), (7, 3), (3, 31)])
def test_general_taps_brute_force(nH, nV):
    """The general-taps extension against the 2-D window definition; with
    symmetric taps it equals FIRFilter2 on the corresponding lobe."""
    rng = np.random.default_rng(nH * 7 + nV)
    th, tv = rng.standard_normal(nH), rng.standard_normal(nV)
    th, tv = (th / np.abs(th).sum()).astype(np.float32), (tv / np.abs(tv).sum()).astype(np.float32)
    img = synth.uniform_image(3, 50, 60)
    rH, rV = nH // 2, nV // 2
    f = img.astype(np.float64)
    acc = np.zeros((50 - 2 * rV, 60 - 2 * rH))
    for i in range(nV):
        for j in range(nH):
            acc += float(tv[i]) * float(th[j]) * f[i:i + 50 - 2 * rV, j:j + 60 - 2 * rH]
    out = np.full(img.shape, np.nan, np.float32)
    oracle.fir2_taps(th, tv, img, out)
    np.testing.assert_allclose(out[rV:rV + acc.shape[0], rH:rH + acc.shape[1]], acc, rtol=1e-5, atol=1e-4)
    assert np.isnan(out[:rV]).all() and np.isnan(out[:, :rH]).all()
    lobe = random_lobe(rng, 9)
    sym = oracle.full_kernel(lobe).astype(np.float32)
    a = np.zeros(img.shape, np.float32)
    b = np.zeros(img.shape, np.float32)
    oracle.fir2(lobe, lobe, img, a)
    oracle.fir2_taps(sym, sym, img, b)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-4)


def _closed_form_failures(method_out, sigma):
    """Closed-form responses of the six methods to polynomial images, for the
    (H, V) lobe pairs of FIRGaussianConvolver.h:181-241.  ``method_out(m,
    img)`` returns method m's output (NaN outside the written region).  With
    k_l the full kernel of lobe l: gain1 = sum_i k1[i] i (ramp response),
    quad2 = sum_i k2[i] (i - c)^2, dc2 = sum_i k2[i], m2 = sum_i k0[i] (i - c)^2.
    Returns the list of checks that fail."""
    l0, l1, l2 = oracle.gauss_lobes(sigma, "f64")
    k0, k1, k2 = (oracle.full_kernel(l) for l in (l0, l1, l2))
    gain1 = float(np.dot(k1, np.arange(k1.size)))
    c0, c2 = (k0.size - 1) // 2, (k2.size - 1) // 2
    quad2 = float(np.dot(k2, (np.arange(k2.size) - c2) ** 2.0))
    dc2 = float(k2.sum())
    m2 = float(np.dot(k0, (np.arange(k0.size) - c0) ** 2.0))
    nrow, ncol = 72, 80
    yc = np.arange(nrow, dtype=np.float64)[:, None] - nrow / 2 + np.zeros((1, ncol))
    xc = np.arange(ncol, dtype=np.float64)[None, :] - ncol / 2 + np.zeros((nrow, 1))
    cases = [
        # image, method, expected output at the same (centre) pixel
        ("y^2 diffVV", yc ** 2, "diffVV", quad2 + dc2 * yc ** 2),
        ("y^2 diffHH", yc ** 2, "diffHH", dc2 * (yc ** 2 + m2)),
        ("x^2 diffHH", xc ** 2, "diffHH", quad2 + dc2 * xc ** 2),
        ("x^2 diffVV", xc ** 2, "diffVV", dc2 * (xc ** 2 + m2)),
        ("xy diffHV", xc * yc, "diffHV", np.full((nrow, ncol), gain1 * gain1)),
        ("xy diffH", xc * yc, "diffH", gain1 * yc),
        ("xy diffV", xc * yc, "diffV", gain1 * xc),
        ("xy smooth", xc * yc, "smooth", xc * yc),
    ]
    bad = []
    for name, img, m, want in cases:
        got = method_out(m, img.astype(np.float32)).astype(np.float64)
        valid = ~np.isnan(got)
        assert valid.sum() > 100, name
        scale = max(1.0, np.abs(want[valid]).max())
        if np.abs(got[valid] - want[valid]).max() > 1e-6 * scale:
            bad.append(name)
    return bad


@pytest.mark.parametrize("sigma", [1.0, 1.5, 2.0, 2.5, 3.0])
def test_second_and_mixed_derivative_closed_forms(sigma):
    """diffVV on a y^2 image, diffHV on an x*y image (and their transposed
    counterparts) pin the (H, V) lobe pairs of diffHV and diffVV; the
    oracle meets every check."""
    lobes = oracle.gauss_lobes(sigma, "f64")

    def method_out(m, img):
        out = np.full(img.shape, np.nan, np.float32)
        return oracle.gauss(m, sigma, img, out, True, "f64", lobes=lobes)

    assert _closed_form_failures(method_out, sigma) == []


@pytest.mark.parametrize("planted,caught", [
    ({"diffVV": (2, 0)}, "y^2 diffVV"),        # diffVV with its lobes swapped
    ({"diffHV": (1, 0)}, "xy diffHV"),         # diffHV with lobe 0 vertically
    ({"diffHV": (0, 1)}, "xy diffHV"),
    ({"diffHH": (0, 2)}, "x^2 diffHH"),
    ({"diffH": (0, 1)}, "xy diffH"),
])
def test_planted_lobe_pair_mistakes_fail(planted, caught):
    """The same checks catch a method wired to the wrong (H, V) lobe pair."""
    sigma = 2.0
    lobes = oracle.gauss_lobes(sigma, "f64")
    pairs = dict(oracle.METHODS, **planted)

    def method_out(m, img):
        h, v = pairs[m]
        out = np.full(img.shape, np.nan, np.float32)
        return oracle.fir2(lobes[h], lobes[v], img, out, True, "f64")

    assert caught in _closed_form_failures(method_out, sigma)
