"""Pins the oracle's ICIA error moment (ICIA.h:61-112, 352-369; vec.h:1285-1290,
1418-1422; array.h:101-118) with closed forms.  On images whose gradients
are constant on rectangles (the valid regions of diffH and diffV of a ramp,
zero elsewhere as in the reference's freshly allocated gradient images),
every moment is a sum of c * s^(p+q) * (sum of x^p over a column range) *
(sum of y^q over a row range), and those power sums have closed forms
(Faulhaber).  A swapped u / v, the wrong normalisation s or a transposed
triangle order fails them."""
import numpy as np
import pytest

import oracle


def _psum(lo, hi, p):
    """sum_{k=lo}^{hi-1} k^p in closed form (p <= 4)."""
    def F(n):                                    # sum_{k=0}^{n-1} k^p
        n = float(n)
        m = n - 1
        if p == 0:
            return n
        if p == 1:
            return m * n / 2
        if p == 2:
            return m * n * (2 * m + 1) / 6
        if p == 3:
            return (m * n / 2) ** 2
        return m * n * (2 * m + 1) * (3 * m * m + 3 * m - 1) / 30
    return F(hi) - F(lo) if hi > lo else 0.0


def _entries(D):
    """a_i as a list of (coefficient name, power of u, power of v):
    a = (eH u, eH v, eH, eV u, eV v, eV[, -eH u^2 - eV v u, -eH u v - eV v^2])."""
    a = [[("H", 1, 0)], [("H", 0, 1)], [("H", 0, 0)], [("V", 1, 0)], [("V", 0, 1)], [("V", 0, 0)]]
    if D == 8:
        a.append([("H", 2, 0, -1.0), ("V", 1, 1, -1.0)])
        a.append([("H", 1, 1, -1.0), ("V", 0, 2, -1.0)])
    return a


def closed_form(D, nrow, ncol, cH, rectH, cV, rectV):
    """Moments for eH = cH on rectH, eV = cV on rectV (rect = (r0, r1, c0, c1),
    rows [r0, r1), columns [c0, c1)), zero elsewhere."""
    s = float(np.float32(1.0) / np.float32(max(nrow, ncol)))
    coef = {"H": (cH, rectH), "V": (cV, rectV)}
    terms = _entries(D)
    out = []
    for i in range(D):
        for j in range(i, D):
            tot = 0.0
            for ti in terms[i]:
                for tj in terms[j]:
                    wi = ti[3] if len(ti) > 3 else 1.0
                    wj = tj[3] if len(tj) > 3 else 1.0
                    ci, ri = coef[ti[0]]
                    cj, rj = coef[tj[0]]
                    # the product is nonzero on the intersection of both rectangles
                    r0, r1 = max(ri[0], rj[0]), min(ri[1], rj[1])
                    c0, c1 = max(ri[2], rj[2]), min(ri[3], rj[3])
                    pu, pv = ti[1] + tj[1], ti[2] + tj[2]
                    tot += wi * wj * ci * cj * s ** (pu + pv) * _psum(c0, c1, pu) * _psum(r0, r1, pv)
            out.append(tot)
    return np.array(out)


def _gradients(nrow, ncol, cH, rectH, cV, rectV):
    eH = np.zeros((nrow, ncol), np.float32)
    eV = np.zeros((nrow, ncol), np.float32)
    eH[rectH[0]:rectH[1], rectH[2]:rectH[3]] = cH
    eV[rectV[0]:rectV[1], rectV[2]:rectV[3]] = cV
    return eH, eV


CASES = [
    (40, 70, 1.5, (4, 36, 2, 68), -0.75, (2, 38, 4, 66)),      # a ramp's diffH / diffV regions
    (90, 60, 2.0, (8, 82, 16, 44), 0.5, (16, 74, 8, 52)),       # taller than wide: s = 1 / nrow
    (33, 33, -1.0, (0, 33, 0, 33), 3.0, (5, 20, 10, 30)),
]


@pytest.mark.parametrize("model", ["affine", "projective"])
@pytest.mark.parametrize("case", CASES)
def test_icia_moment_closed_forms(model, case):
    nrow, ncol, cH, rH, cV, rV = case
    eH, eV = _gradients(nrow, ncol, cH, rH, cV, rV)
    got, absm = oracle.icia_moment(eH, eV, model)
    want = closed_form(oracle.ICIA_MODELS[model], nrow, ncol, cH, rH, cV, rV)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    assert (absm >= np.abs(got) - 1e-9).all()


def test_icia_planted_mistakes_fail():
    """u and v swapped, or s = 1 / min(nrow, ncol): the closed forms differ."""
    nrow, ncol, cH, rH, cV, rV = CASES[1]
    want = closed_form(6, nrow, ncol, cH, rH, cV, rV)
    eH, eV = _gradients(nrow, ncol, cH, rH, cV, rV)
    # u / v swapped = the transposed images with the gradients exchanged roles
    swapped, _ = oracle.icia_moment(eH.T.copy(), eV.T.copy(), "affine")
    assert np.abs(swapped - want).max() > 1e-3 * np.abs(want).max()
    # the wrong normalisation: the moments of an image padded to a square
    # (s = 1 / max) differ from s = 1 / min by the factors s^(p+q)
    s_min = 1.0 / min(nrow, ncol)
    s_max = 1.0 / max(nrow, ncol)
    assert abs(s_min - s_max) > 0
    wrong = closed_form(6, nrow, ncol, cH, rH, cV, rV) * 0
    terms_scale = []
    for i in range(6):
        for j in range(i, 6):
            pw = (i % 3 < 2) + (j % 3 < 2)
            terms_scale.append((s_min / s_max) ** pw)
    wrong = want * np.array(terms_scale)
    assert np.abs(wrong - want).max() > 1e-3 * np.abs(want).max()


def test_icia_triangle_order():
    """Entry k of the moment is a_i a_j for the k-th (i, j), i <= j, row by
    row (array.h:108: out[D i + j - i (i + 1) / 2]): one nonzero a_i per
    image isolates entries."""
    nrow, ncol = 10, 20
    eH = np.zeros((nrow, ncol), np.float32)
    eV = np.zeros((nrow, ncol), np.float32)
    eH[0, 0] = 2.0          # u = v = 0: only a_2 = eH is nonzero
    got, _ = oracle.icia_moment(eH, eV, "affine")
    k = 6 * 2 + 2 - 2 * 3 // 2
    want = np.zeros(21)
    want[k] = 4.0
    np.testing.assert_array_equal(got, want)
    eV[0, 0] = 3.0          # a_5 = eV too: entries (2,2), (2,5), (5,5)
    got, _ = oracle.icia_moment(eH, eV, "affine")
    want[6 * 2 + 5 - 3] = 6.0
    want[6 * 5 + 5 - 15] = 9.0
    np.testing.assert_array_equal(got, want)
