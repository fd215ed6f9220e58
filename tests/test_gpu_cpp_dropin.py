"""The C++ drop-in (include/TU/cu/*.h) driven the way the reference's demos
drive the reference (Testprog/gaussTest/main.cu, Testprog/filterTest/main.cu):
tests/cpp/dropin_demo filters a PGM through cu::Array2, FIRGaussianConvolver2
and FIRFilter2 on the GPU; its outputs are compared with the oracle here."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests._util import assert_f32_parity, assert_u8_parity
from tucuda_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "dropin_demo")


def read_pnm(path):
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(b"\n", 3)
    magic = parts[0]
    w, h = map(int, parts[1].split())
    if magic == b"P5":
        return np.frombuffer(parts[3], np.uint8, w * h).reshape(h, w)
    assert magic == b"Pf"
    return np.frombuffer(parts[3], "<f4", w * h).reshape(h, w)[::-1]


def coefficients(sigma, lobe_size):
    """Testprog/filterTest/main.cu:17-47 as the demo evaluates it."""
    s = np.float32(sigma)
    c = np.zeros(lobe_size + 1, np.float32)
    for i in range(lobe_size + 1):
        dx = np.float32(lobe_size - i) / s
        c[i] = np.float32(np.exp(-0.5 * np.float64(dx) * np.float64(dx)))
    total = np.float32(c[lobe_size])
    for i in range(lobe_size):
        total = np.float32(total + np.float32(2.0) * c[i])
    return (c / total).astype(np.float32)


def test_dropin_demo(dev, tmp_path):
    assert os.path.exists(DEMO), "build the C++ demo first (__graft_entry__.build())"
    sigma, lobe_size = 2.0, 16
    img = synth.u8_image(21, 150, 260)
    pgm = tmp_path / "in.pgm"
    with open(pgm, "wb") as f:
        f.write(b"P5\n260 150\n255\n" + img.tobytes())
    out = subprocess.run([DEMO, str(pgm), str(sigma), str(lobe_size), str(tmp_path), "20"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    info = json.loads(out.stdout.strip().splitlines()[-1])
    assert info["too_large"] == "FIRFilter2<T, BLOCK_TRAITS>::initialize: too large lobe size!"
    assert info["unsupported"] == "FIRFilter2::convolve: unsupported horizontal lobe size[6]!"
    assert info["offsetH"] == 6

    lobes = oracle.gauss_lobes(sigma)
    ref_u8 = np.zeros_like(img)
    ref_f = np.zeros(img.shape, np.float32)
    oracle.gauss("smooth", sigma, img, ref_u8, lobes=lobes)
    oracle.gauss("smooth", sigma, img, ref_f, lobes=lobes)
    assert_u8_parity(read_pnm(tmp_path / "smooth_u8.pgm"), ref_u8, ref_f, "drop-in smooth u8")
    for m in oracle.METHOD_ORDER:
        ref = np.zeros(img.shape, np.float32)
        oracle.gauss(m, sigma, img, ref, lobes=lobes)
        assert_f32_parity(read_pnm(tmp_path / f"{m}_f32.pfm"), ref, f"drop-in {m}")
        assert_f32_parity(read_pnm(tmp_path / f"fused_{m}.pfm"), ref, f"drop-in fused {m}")
    coeff = coefficients(sigma, lobe_size)
    ref = np.zeros(img.shape, np.float32)
    oracle.fir2(coeff, coeff, img.astype(np.float32), ref)
    assert_f32_parity(read_pnm(tmp_path / "fir_f32.pfm"), ref, "drop-in FIRFilter2")

    # the pyramid step: subsample(identity1x1) of the drop-in's own smooth
    # output is its exact decimation; gaussian5x5 of the float upload equals
    # the oracle bit for bit (no contraction on either side)
    smooth = read_pnm(tmp_path / "smooth_f32.pfm")
    half = read_pnm(tmp_path / "half_smooth.pfm")
    np.testing.assert_array_equal(half, smooth[0:150:2, 0:260:2])
    assert_f32_parity(read_pnm(tmp_path / "half_smooth_fused.pfm"), half, "drop-in smoothSubsample")
    ref = np.zeros((75, 130), np.float32)
    oracle.subsample(img.astype(np.float32), ref, "gaussian5x5")
    np.testing.assert_array_equal(read_pnm(tmp_path / "gaussian5x5_half.pfm"), ref)


def test_dropin_slice_and_make_range(dev, tmp_path):
    """cu::slice(it, 10, nrow-21, 7, ncol-21) and cu::make_range(it + 3,
    nrow - 6) as inputs and outputs: each result equals the oracle on the
    sub-image, placed in the same window, and nothing outside it is
    written."""
    assert os.path.exists(DEMO), "build the C++ demo first (__graft_entry__.build())"
    sigma = 1.5
    img = synth.u8_image(22, 120, 200)
    pgm = tmp_path / "in.pgm"
    with open(pgm, "wb") as f:
        f.write(b"P5\n200 120\n255\n" + img.tobytes())
    out = subprocess.run([DEMO, str(pgm), str(sigma), "8", str(tmp_path), "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    f = img.astype(np.float32)
    lobes = oracle.gauss_lobes(sigma)
    got = read_pnm(tmp_path / "slice_diffH.pfm")
    sub = np.ascontiguousarray(f[10:10 + 120 - 21, 7:7 + 200 - 21])
    ref = np.zeros(sub.shape, np.float32)
    oracle.gauss("diffH", sigma, sub, ref, lobes=lobes)
    assert_f32_parity(got[10:10 + sub.shape[0], 7:7 + sub.shape[1]], ref, "slice diffH")
    outside = np.ones(got.shape, bool)
    outside[10:10 + sub.shape[0], 7:7 + sub.shape[1]] = False
    assert (got[outside] == 0).all()
    got = read_pnm(tmp_path / "range_smooth.pfm")
    rows = np.ascontiguousarray(f[3:120 - 3])
    ref = np.zeros(rows.shape, np.float32)
    oracle.gauss("smooth", sigma, rows, ref, lobes=lobes)
    assert_f32_parity(got[3:120 - 3], ref, "make_range smooth")
    assert (got[:3] == 0).all() and (got[120 - 3:] == 0).all()
