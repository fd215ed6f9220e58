"""Pins the oracle's cu::subsample restatement (TU/cu/algorithm.h:183-242,
operators of TU/cu/functional.h:58-256) without the reference, which cannot
be built here and holds no recorded outputs for it:

* identity1x1 is plain decimation: ``img[0:2*(n//2):2, 0:2*(m//2):2]``,
  exact for every element type;
* on polynomial images each linear operator has a closed form (a ramp's
  slope, a parabola's curvature, the mixed derivative of x*y, the Hessian
  determinant of x^2+y^2 and of x*y, the binomial 5x5's unit variance);
* each operator's border band is zero exactly where its stencil leaves the
  image, with the guard widths of functional.h (1x3 operators guard columns
  only, 3x1 rows only, 3x3 both by one pixel, gaussian5x5 both by two);
* on random data every operator agrees with an independent float64
  formulation (2-D correlation masks, abs / products for the non-linear
  ones) to FP32 rounding;
* a planted mistake (a transposed mask) is caught by the same comparison.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle

OPS = oracle.SUBSAMPLE_OPS
# guard widths (rows, columns): functional.h's (OperatorSizeY-1)/2, (OperatorSizeX-1)/2
GUARD = {"identity1x1": (0, 0), "diffH1x3": (0, 1), "diffV3x1": (1, 0), "diffHH1x3": (0, 1),
         "diffVV3x1": (1, 0), "diffHV3x3": (1, 1), "sobelH3x3": (1, 1), "sobelV3x3": (1, 1),
         "sobelAbs3x3": (1, 1), "laplacian3x3": (1, 1), "det3x3": (1, 1), "gaussian5x5": (2, 2)}


def _run(img, op, out_dtype=np.float32):
    nrow, ncol = img.shape
    out = np.full((max(nrow // 2, 1), max(ncol // 2, 1)), 7, dtype=out_dtype)
    return oracle.subsample(np.ascontiguousarray(img), out, op)


def _poly(nrow, ncol, f):
    y, x = np.mgrid[0:nrow, 0:ncol].astype(np.float64)
    return f(y, x).astype(np.float32)


@pytest.mark.parametrize("dt", [np.uint8, np.float32, np.float16, np.uint16])
@pytest.mark.parametrize("shape", [(6, 8), (7, 9), (2, 2), (33, 65)])
def test_identity_is_decimation(dt, shape):
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, size=shape).astype(dt)
    if dt == np.uint16:   # bfloat16 bit patterns of small integers
        img = (np.frombuffer(img.astype(np.float32).tobytes(), dtype=np.uint32) >> 16).astype(np.uint16).reshape(shape)
    got = _run(img, "identity1x1")
    n, m = shape[0] // 2 * 2, shape[1] // 2 * 2
    want = img[0:n:2, 0:m:2]
    if dt == np.uint16:
        want = np.frombuffer((want.astype(np.uint32) << 16).tobytes(), dtype=np.float32).reshape(want.shape)
    np.testing.assert_array_equal(got, want.astype(np.float32))


# (image, operator) -> closed-form interior value as a function of (Y, X)
CLOSED = [
    (lambda y, x: 3 * x - 2 * y + 5, "diffH1x3", lambda Y, X: 3.0),
    (lambda y, x: 3 * x - 2 * y + 5, "diffV3x1", lambda Y, X: -2.0),
    (lambda y, x: 3 * x - 2 * y + 5, "diffHH1x3", lambda Y, X: 0.0),
    (lambda y, x: 3 * x - 2 * y + 5, "diffHV3x3", lambda Y, X: 0.0),
    (lambda y, x: 3 * x - 2 * y + 5, "sobelH3x3", lambda Y, X: 3.0),
    (lambda y, x: 3 * x - 2 * y + 5, "sobelV3x3", lambda Y, X: -2.0),
    (lambda y, x: 3 * x - 2 * y + 5, "sobelAbs3x3", lambda Y, X: 5.0),
    (lambda y, x: 3 * x - 2 * y + 5, "laplacian3x3", lambda Y, X: 0.0),
    (lambda y, x: 3 * x - 2 * y + 5, "gaussian5x5", lambda Y, X: 3 * X - 2 * Y + 5),
    (lambda y, x: x * x, "diffHH1x3", lambda Y, X: 2.0),
    (lambda y, x: x * x, "diffVV3x1", lambda Y, X: 0.0),
    (lambda y, x: x * x, "laplacian3x3", lambda Y, X: 2.0),
    (lambda y, x: x * x, "diffH1x3", lambda Y, X: 2.0 * X),
    (lambda y, x: y * y, "diffVV3x1", lambda Y, X: 2.0),
    (lambda y, x: y * y, "diffV3x1", lambda Y, X: 2.0 * Y),
    (lambda y, x: x * x, "gaussian5x5", lambda Y, X: X * X + 1.0),     # sum w_i i^2 = 1
    (lambda y, x: y * y, "gaussian5x5", lambda Y, X: Y * Y + 1.0),
    (lambda y, x: x * y, "diffHV3x3", lambda Y, X: 1.0),
    (lambda y, x: x * y, "det3x3", lambda Y, X: -1.0),
    (lambda y, x: x * y, "gaussian5x5", lambda Y, X: X * Y),
    (lambda y, x: x * x + y * y, "det3x3", lambda Y, X: 4.0),
    (lambda y, x: x * x + y * y, "laplacian3x3", lambda Y, X: 4.0),
    (lambda y, x: x * x - y * y, "det3x3", lambda Y, X: -4.0),
]


@pytest.mark.parametrize("case", range(len(CLOSED)))
def test_closed_forms_and_border(case):
    f, op, g = CLOSED[case]
    nrow, ncol = 23, 30
    got = _run(_poly(nrow, ncol, f), op)
    gy, gx = GUARD[op]
    for yo in range(nrow // 2):
        for xo in range(ncol // 2):
            Y, X = 2 * yo, 2 * xo
            inside = gy <= Y < nrow - gy and gx <= X < ncol - gx
            want = g(Y, X) if inside else 0.0
            assert got[yo, xo] == pytest.approx(want, abs=1e-4), (op, yo, xo)


def _masks():
    z = np.zeros((5, 5))

    def m(entries):
        a = z.copy()
        for (dy, dx), w in entries.items():
            a[2 + dy, 2 + dx] = w
        return a
    d = {
        "diffH1x3": m({(0, 1): 0.5, (0, -1): -0.5}),
        "diffV3x1": m({(1, 0): 0.5, (-1, 0): -0.5}),
        "diffHH1x3": m({(0, -1): 1, (0, 0): -2, (0, 1): 1}),
        "diffVV3x1": m({(-1, 0): 1, (0, 0): -2, (1, 0): 1}),
        "diffHV3x3": m({(-1, -1): 0.25, (-1, 1): -0.25, (1, -1): -0.25, (1, 1): 0.25}),
        "sobelH3x3": m({(-1, 1): 0.125, (-1, -1): -0.125, (1, 1): 0.125, (1, -1): -0.125,
                        (0, 1): 0.25, (0, -1): -0.25}),
        "laplacian3x3": m({(0, -1): 1, (0, 1): 1, (-1, 0): 1, (1, 0): 1, (0, 0): -4}),
    }
    d["sobelV3x3"] = d["sobelH3x3"].T.copy()
    b = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
    d["gaussian5x5"] = np.outer(b, b)
    d["identity1x1"] = m({(0, 0): 1})
    return d


def _corr(img, mask):
    """Full-resolution 2-D correlation in float64 (zero outside), centred."""
    n, m = img.shape
    p = np.zeros((n + 4, m + 4))
    p[2:-2, 2:-2] = img
    out = np.zeros((n, m))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            w = mask[2 + dy, 2 + dx]
            if w:
                out += w * p[2 + dy:2 + dy + n, 2 + dx:2 + dx + m]
    return out


def independent(img, op, masks=None):
    """float64 formulation of each operator, decimated and guarded."""
    masks = masks or _masks()
    img = img.astype(np.float64)
    if op == "sobelAbs3x3":
        full = np.abs(_corr(img, masks["sobelH3x3"])) + np.abs(_corr(img, masks["sobelV3x3"]))
    elif op == "det3x3":
        full = _corr(img, masks["diffHH1x3"]) * _corr(img, masks["diffVV3x1"]) - _corr(img, masks["diffHV3x3"]) ** 2
    else:
        full = _corr(img, masks[op])
    n, m = img.shape
    gy, gx = GUARD[op]
    ys = np.arange(n)[:, None]
    xs = np.arange(m)[None, :]
    full = np.where((ys >= gy) & (ys < n - gy) & (xs >= gx) & (xs < m - gx), full, 0.0)
    return full[0:n // 2 * 2:2, 0:m // 2 * 2:2]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("shape", [(40, 36), (37, 51)])
def test_random_against_float64(op, shape):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 255, size=shape).astype(np.float32)
    got = _run(img, op).astype(np.float64)
    want = independent(img, op)
    scale = 255.0 ** 2 if op == "det3x3" else 255.0
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * scale)


def test_planted_transpose_is_caught():
    masks = _masks()
    masks["sobelH3x3"] = masks["sobelH3x3"].T.copy()      # the mistake
    rng = np.random.default_rng(6)
    img = rng.uniform(0, 255, size=(30, 30)).astype(np.float32)
    got = _run(img, "sobelH3x3").astype(np.float64)
    assert not np.allclose(got, independent(img, "sobelH3x3", masks), rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("shape", [(1, 9), (9, 1), (1, 1), (1, 2)])
def test_too_small_is_noop(shape):
    img = np.ones(shape, dtype=np.float32)
    out = np.full((8, 8), 7, dtype=np.float32)
    oracle.subsample(img, out, "identity1x1")
    assert (out == 7).all()


def test_u8_output_saturates_and_truncates():
    img = _poly(8, 8, lambda y, x: 100.0 * x)               # diffH = 100; negative ramp -> -100
    assert (_run(img, "diffH1x3", np.uint8)[:, 1:] == 100).all()
    img = _poly(8, 8, lambda y, x: -100.0 * x + 0.7 * 2 * x)
    assert (_run(img, "diffH1x3", np.uint8) == 0).all()
    img = _poly(8, 8, lambda y, x: 600.0 * x)
    assert (_run(img, "diffH1x3", np.uint8)[:, 1:] == 255).all()
    img = _poly(8, 8, lambda y, x: 3.75 * x)
    assert (_run(img, "diffH1x3", np.uint8)[:, 1:] == 3).all()
