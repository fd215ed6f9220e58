"""Pins the oracle's Gaussian lobes (FIRGaussianConvolver.h:123-170,
FIRGaussianConvolver.cu:52-98) with checks that do not retype its code:
closed-form truncation radii, the normalisation identities each lobe is
built to satisfy (unit DC gain, the ramp and quadratic responses), and the
size-class table derived independently of the oracle."""
import math

import numpy as np
import pytest

import oracle
from tests._util import full_taps

SIGMAS = np.round(np.arange(0.55, 4.6, 0.0125), 6)


def _classes(siz: int, even: bool) -> int:
    if even:
        return 3 if siz <= 2 else 5 if siz <= 5 else 9 if siz <= 9 else 17
    return 2 if siz <= 1 else 4 if siz <= 4 else 8 if siz <= 8 else 16


def _closed_form_sizes(sigma: float):
    """Lobe lengths from the closed forms of the sampled functions:
    g(i) = exp(-i^2 / 2 sigma^2) peaks at i = 0, so it stays above 1 % of its
    peak for i < sigma * sqrt(2 ln 100); |g'(i)| = (i/sigma) g(i) and
    |g''(i)| = |i^2/sigma^2 - 1| g(i) are thresholded against their largest
    sample.  Returns None when a threshold crossing is too close to call."""
    i = np.arange(17, dtype=np.float64)
    g = np.exp(-0.5 * (i / sigma) ** 2)
    out = []
    for vals, even in ((g, True), ((i / sigma) * g, False), (np.abs((i / sigma) ** 2 - 1) * g, True)):
        th = 0.01 * vals.max()
        rel = np.abs(vals - th) / th
        if (rel < 1e-4).any():
            return None
        above = np.nonzero(vals > th)[0]
        out.append(_classes(int(above[-1]) + 1, even))
    return tuple(out)


def test_lobe0_radius_closed_form():
    """siz0 - 1 = floor(sigma * sqrt(2 ln 100)) away from ties."""
    for sigma in SIGMAS:
        r = sigma * math.sqrt(2 * math.log(100))
        if abs(r - round(r)) < 1e-3:
            continue
        siz = min(int(math.floor(r)) + 1, 17)
        n0 = oracle.gauss_lobes(float(sigma))[0].size
        assert n0 == _classes(siz, True), sigma


def test_sizes_match_closed_forms():
    checked = 0
    for sigma in SIGMAS:
        want = _closed_form_sizes(float(sigma))
        if want is None:
            continue
        got = tuple(x.size for x in oracle.gauss_lobes(float(sigma)))
        assert got == want, (sigma, got, want)
        assert tuple(x.size for x in oracle.gauss_lobes(float(sigma), "f64")) == want
        checked += 1
    assert checked > 250


@pytest.mark.parametrize("sigma,sizes", [
    (0.8, (5, 4, 5)), (1.0, (5, 4, 5)), (1.2, (5, 8, 5)), (1.5, (5, 8, 9)),
    (2.0, (9, 8, 9)), (2.3, (9, 16, 9)), (2.5, (9, 16, 17)), (3.0, (17, 16, 17)), (4.0, (17, 16, 17)),
])
def test_size_table(sigma, sizes):
    """The sigma -> (l0, l1, l2) length classes of the configurations."""
    assert tuple(x.size for x in oracle.gauss_lobes(sigma)) == sizes


def _samples(sigma):
    i = np.arange(17, dtype=np.float64)
    g = np.exp(-0.5 * (i / sigma) ** 2)
    return g, -(i / sigma) * g, ((i / sigma) ** 2 - 1) * g


@pytest.mark.parametrize("sigma", [0.8, 1.0, 1.3, 1.7, 2.0, 2.5, 3.0, 3.7, 4.0])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_normalisation_identities(sigma, precision):
    l0, l1, l2 = oracle.gauss_lobes(sigma, precision)
    tol = 2e-6 if precision == "f32" else 1e-13
    k0, k1, k2 = (oracle.full_kernel(l) for l in (l0, l1, l2))
    # lobe0: unit DC gain, symmetric, peak at the centre (last entry).
    assert abs(k0.sum() - 1.0) < tol
    assert np.argmax(l0) == l0.size - 1
    # lobe1: antisymmetric; the ramp x[j] = j gives 1 + (2S g1(S)) / sum_{i<S} 2 i g1(i),
    # the normalisation over i < S applied to taps 1..S.
    S = l1.size
    _, g1, g2 = _samples(sigma)
    j = np.arange(k1.size, dtype=np.float64)
    ramp = float(np.dot(k1, j))
    want = 1.0 + 2 * S * abs(g1[S]) / sum(2 * i * abs(g1[i]) for i in range(S))
    assert abs(ramp - want) < 10 * tol, (ramp, want)
    assert abs(np.dot(k1, np.ones_like(j))) < tol               # no DC response
    # lobe2: quadratic response +-2 (normalised by |sum i^2 g2(i)|), DC in closed form.
    S2 = l2.size
    den = sum(i * i * g2[i] for i in range(1, S2))
    c = (k2.size - 1) // 2
    quad = float(np.dot(k2, (np.arange(k2.size) - c) ** 2.0))
    assert abs(quad - 2.0 * np.sign(den)) < 20 * tol, quad
    dc = (g2[0] + 2 * g2[1:S2].sum()) / abs(den)
    assert abs(k2.sum() - dc) < 20 * tol


def test_f32_and_f64_lobes_agree():
    for sigma in (0.8, 1.5, 2.0, 3.0):
        for a, b in zip(oracle.gauss_lobes(sigma), oracle.gauss_lobes(sigma, "f64")):
            assert a.size == b.size
            np.testing.assert_allclose(a, b, rtol=3e-6, atol=1e-9)


def test_lobe_size_degenerate():
    """All-zero samples: nothing exceeds the threshold, the search index wraps
    and the largest class is returned (FIRGaussianConvolver.cu:69-97)."""
    assert oracle.lobe_size(np.zeros(17), True) == 17
    assert oracle.lobe_size(np.zeros(17), False) == 16
    e = np.zeros(17)
    e[0] = 1.0
    assert oracle.lobe_size(e, True) == 3
    assert oracle.lobe_size(e, False) == 2


# ---------------------------------------------------------------------------
# The ceil(4 sigma) radius rule (batched and served workloads): radius fixed
# at r = ceil(4 sigma) clamped to [1, 16], the reference's normalisations.

CEIL4_SIGMAS = [0.8, 0.9, 1.0, 1.13, 1.3, 1.6, 2.0, 2.2, 2.45, 2.8, 3.0, 3.3, 3.7, 3.99, 4.0]


def _ceil4_radius(sigma: float, precision: str) -> int:
    s = float(np.float32(sigma)) if precision == "f32" else sigma
    return max(1, min(16, math.ceil(4 * s)))


@pytest.mark.parametrize("sigma", CEIL4_SIGMAS)
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_ceil4_sizes_and_sample_shape(sigma, precision):
    """Lengths r + 1, r, r + 1; every entry is the sampled Gaussian, its
    first or its second derivative at the offset the entry stands for
    (lobe[0] outermost), up to one positive factor per lobe."""
    l0, l1, l2 = oracle.gauss_lobes(sigma, precision, "ceil4")
    r = _ceil4_radius(sigma, precision)
    assert (l0.size, l1.size, l2.size) == (r + 1, r, r + 1)
    s = float(np.float32(sigma)) if precision == "f32" else sigma
    tol = 5e-6 if precision == "f32" else 1e-12
    off0 = np.arange(r, -1, -1, dtype=np.float64)          # offsets of lobe0 / lobe2 entries
    off1 = np.arange(r, 0, -1, dtype=np.float64)           # offsets of lobe1 entries
    g = lambda o: np.exp(-0.5 * (o / s) ** 2)
    want0 = g(off0)
    want1 = -(off1 / s) * g(off1)
    want2 = ((off0 / s) ** 2 - 1) * g(off0)
    for got, want in ((l0, want0), (l1, want1), (l2, want2)):
        got = got.astype(np.float64)
        k = np.argmax(np.abs(want))
        factor = got[k] / want[k]
        assert factor > 0
        np.testing.assert_allclose(got, factor * want, rtol=tol, atol=tol * np.abs(got).max())


@pytest.mark.parametrize("sigma", CEIL4_SIGMAS)
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_ceil4_normalisation_identities(sigma, precision):
    """The identities the reference's normalisation builds in
    (FIRGaussianConvolver.h:141-163), at the fixed radius: unit DC gain of
    lobe 0; lobe 1's ramp response 1 + 2 r |g1(r)| / sum_{i<r} 2 i |g1(i)|
    (taps at offsets r .. 1, normalised over 0 .. r-1) and no DC response;
    lobe 2's quadratic response +-2 and its DC in closed form."""
    l0, l1, l2 = oracle.gauss_lobes(sigma, precision, "ceil4")
    r = l1.size
    s = float(np.float32(sigma)) if precision == "f32" else sigma
    tol = 3e-6 if precision == "f32" else 1e-13
    # lobe roles, not lengths, fix the parity here (r + 1 is even for odd r)
    k0, k1, k2 = full_taps(l0, True), full_taps(l1, False), full_taps(l2, True)
    i = np.arange(17, dtype=np.float64)
    g0 = np.exp(-0.5 * (i / s) ** 2)
    g1 = -(i / s) * g0
    g2 = ((i / s) ** 2 - 1) * g0
    assert abs(k0.sum() - 1.0) < tol
    ramp = float(np.dot(k1, np.arange(k1.size, dtype=np.float64)))
    want = 1.0 + 2 * r * abs(g1[r]) / sum(2 * j * abs(g1[j]) for j in range(r))
    assert abs(ramp - want) < 20 * tol, (ramp, want)
    assert abs(k1.sum()) < tol
    den = sum(j * j * g2[j] for j in range(1, r + 1))
    quad = float(np.dot(k2, (np.arange(k2.size) - r) ** 2.0))
    assert abs(quad - 2.0 * np.sign(den)) < 40 * tol, quad
    dc = (g2[0] + 2 * g2[1:r + 1].sum()) / abs(den)
    assert abs(k2.sum() - dc) < 40 * tol


def test_ceil4_f32_and_f64_agree():
    for sigma in CEIL4_SIGMAS:
        for a, b in zip(oracle.gauss_lobes(sigma, "f32", "ceil4"), oracle.gauss_lobes(sigma, "f64", "ceil4")):
            assert a.size == b.size
            np.testing.assert_allclose(a, b, rtol=3e-6, atol=1e-8)


def test_ceil4_clamps():
    """Radii below 1 and above 16 are clamped."""
    assert [x.size for x in oracle.gauss_lobes(0.1, "f64", "ceil4")] == [2, 1, 2]
    assert [x.size for x in oracle.gauss_lobes(9.0, "f64", "ceil4")] == [17, 16, 17]


def test_planted_ceil4_mistakes_fail():
    """A lobe 1 normalised over offsets 1 .. r (no off-by-one) or a radius
    one short breaks the checks above: the identities pin the rule."""
    sigma, r = 2.2, 9
    s = sigma
    i = np.arange(17, dtype=np.float64)
    g1 = -(i / s) * np.exp(-0.5 * (i / s) ** 2)
    mutant = g1[r:0:-1] / abs(sum(2 * j * g1[j] for j in range(1, r + 1)))     # normalised over 1..r
    k1 = full_taps(mutant, False)
    ramp = float(np.dot(k1, np.arange(k1.size, dtype=np.float64)))
    want = 1.0 + 2 * r * abs(g1[r]) / sum(2 * j * abs(g1[j]) for j in range(r))
    assert abs(ramp - want) > 1e-3
    assert oracle.gauss_lobes(sigma, "f64", "ceil4")[1].size == r != r - 1
