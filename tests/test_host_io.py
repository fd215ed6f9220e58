"""Host image I/O of the C++ drop-in (include/TU/cu/host.h, the stand-in
for TULibs' Image<T>::restore / save that the reference's demos use,
Testprog/gaussTest/main.cu:85-87, 108-109), on the CPU: streams shaped like
the reference's sample inputs -- a P5 stream of several images with TULibs'
EPBM comment headers (as Testprog/ueshiba.epbm holds), a P6 colour image (as
Testprog/src0.ppm) -- and float PFM in both byte orders, grey and colour."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no host C++ compiler")
    exe = tmp_path_factory.mktemp("hostio") / "host_io_check"
    subprocess.run([cxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "host_io_check.cpp")], check=True)
    return str(exe)


def _run(checker, kind, data, tmp_path):
    src, dst = tmp_path / "in.pnm", tmp_path / "out.pnm"
    src.write_bytes(data)
    r = subprocess.run([checker, kind, str(src), str(dst)], capture_output=True, text=True, timeout=60)
    return r, [tuple(float(v) for v in line.split()) for line in r.stdout.split("\n") if line.strip()], dst


EPBM_KEYS = (b"# PixelLength: 1\n# DataType: Char\n# Sign: Unsigned\n# Endian: Big\n"
             b"# PinHoleParameterH11: 1128.72\n# PinHoleParameterF: 1.0\n")


def _p5(img, comments=b""):
    h, w = img.shape
    return b"P5\n" + comments + f"{w} {h}\n255\n".encode() + img.astype(np.uint8).tobytes()


def test_epbm_stream_of_three(checker, tmp_path):
    rng = np.random.default_rng(1)
    imgs = [rng.integers(0, 256, size=s).astype(np.uint8) for s in ((5, 7), (3, 4), (2, 6))]
    r, rows, dst = _run(checker, "u8", b"".join(_p5(i, EPBM_KEYS) for i in imgs), tmp_path)
    assert r.returncode == 0, r.stderr
    assert rows == [(i.shape[1], i.shape[0], float(i.sum())) for i in imgs]
    out = dst.read_bytes()
    assert out == _p5(imgs[-1])                         # saved back as plain P5


def test_p6_colour_to_grey(checker, tmp_path):
    rgb = np.random.default_rng(2).integers(0, 256, size=(4, 5, 3)).astype(np.uint8)
    data = b"P6\n# a comment\n5 4\n255\n" + rgb.tobytes()
    r, rows, _ = _run(checker, "u8", data, tmp_path)
    assert r.returncode == 0, r.stderr
    grey = rgb.astype(np.uint32).sum(axis=2) // 3
    assert rows == [(5.0, 4.0, float(grey.sum()))]


@pytest.mark.parametrize("little", [True, False])
@pytest.mark.parametrize("colour", [False, True])
def test_pfm(checker, tmp_path, little, colour):
    rng = np.random.default_rng(3)
    h, w = 3, 5
    img = rng.uniform(-10, 10, size=(h, w, 3 if colour else 1)).astype(np.float32)
    order = "<" if little else ">"
    body = img[::-1].astype(order + "f4").tobytes()      # bottom row first
    data = (b"PF\n" if colour else b"Pf\n") + f"{w} {h}\n{-1.0 if little else 1.0}\n".encode() + body
    r, rows, dst = _run(checker, "f32", data, tmp_path)
    assert r.returncode == 0, r.stderr
    grey = img.mean(axis=2, dtype=np.float32) if colour else img[..., 0]
    (ww, hh, s), = rows
    assert (ww, hh) == (w, h) and s == pytest.approx(float(grey.astype(np.float64).sum()), rel=1e-5, abs=1e-4)
    out = dst.read_bytes()                               # written back as little-endian grey Pf
    head = f"Pf\n{w} {h}\n-1.0\n".encode()
    assert out.startswith(head)
    back = np.frombuffer(out[len(head):], "<f4").reshape(h, w)[::-1]
    np.testing.assert_allclose(back, grey, rtol=1e-6)


def test_truncated_is_an_error(checker, tmp_path):
    r, _, _ = _run(checker, "u8", b"P5\n4 4\n255\n" + bytes(10), tmp_path)
    assert r.returncode == 1 and "truncated" in r.stderr


def test_unknown_magic_is_an_error(checker, tmp_path):
    r, _, _ = _run(checker, "u8", b"P3\n2 2\n255\n1 2 3 4\n", tmp_path)
    assert r.returncode == 1 and "unsupported" in r.stderr
