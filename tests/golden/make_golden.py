"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

The reference ships no recorded outputs for this path, so the fixtures are
the oracle's FP32-faithful results on seeded inputs (SplitMix64,
tucuda_amd/synth.py), frozen: tests/test_golden.py checks that the oracle
and the host library still reproduce them bit for bit, and
tests/test_gpu_golden.py holds the GPU kernels to the FP32 bar against them.

Contents:
  lobes.json      FIRGaussianConvolver2 lobes (reference rule) for eight sigmas,
                  as float32 hex strings
  gauss.npz       a 64x48 u_char image (seed 11) and a 97x61 float32 image
                  (seed 14); the six methods' float32 outputs (shift = true) at sigma
                  0.8, 1.2, 2.0, 2.5, 3.0 (lobe classes (5,4,5), (5,8,5),
                  (9,8,9), (9,16,17), (17,16,17))
  fir.npz         FIRFilter2 with N(0,1) lobes (each full kernel scaled to
                  unit L1 norm) of four length pairs on the 97x61 float32
                  image, both shifts
  subsample.npz   cu::subsample of a 37x51 float32 image with each operator,
                  and identity1x1 of a u_char image
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle                      # noqa: E402
from tucuda_amd import synth       # noqa: E402

LOBE_SIGMAS = (0.8, 1.0, 1.2, 1.5, 2.0, 2.5, 3.0, 4.0)
GAUSS_SIGMAS = (0.8, 1.2, 2.0, 2.5, 3.0)
FIR_PAIRS = ((3, 5), (2, 9), (16, 17), (8, 4))


def images():
    return {
        "u8_64x48": synth.u8_image(11, 48, 64),
        "f32_97x61": synth.edge_image(14, 61, 97),
    }


def main():
    lobes = []
    for s in LOBE_SIGMAS:
        l0, l1, l2 = oracle.gauss_lobes(s)
        lobes.append({"sigma": s, **{f"lobe{k}": [float(v).hex() for v in l] for k, l in enumerate((l0, l1, l2))}})
    with open(os.path.join(HERE, "lobes.json"), "w") as f:
        json.dump(lobes, f, indent=1)

    imgs = images()
    g = dict(imgs)
    for name, img in imgs.items():
        for s in GAUSS_SIGMAS:
            lb = oracle.gauss_lobes(s)
            for m in oracle.METHOD_ORDER:
                out = np.zeros(img.shape, np.float32)
                oracle.gauss(m, s, img, out, lobes=lb)
                g[f"{name}_s{s}_{m}"] = out
    np.savez_compressed(os.path.join(HERE, "gauss.npz"), **g)

    img = imgs["f32_97x61"]
    fir = {"image": img}
    for k, (nh, nv) in enumerate(FIR_PAIRS):
        z = synth.normal(30 + k, nh + nv)
        # N(0,1) lobes scaled so that each full kernel has unit L1 norm (the
        # FP32 bar's absolute floor assumes a kernel of unit gain)
        lh = (z[:nh] / np.abs(oracle.full_kernel(z[:nh])).sum()).astype(np.float32)
        lv = (z[nh:] / np.abs(oracle.full_kernel(z[nh:])).sum()).astype(np.float32)
        fir[f"p{k}_lobeH"], fir[f"p{k}_lobeV"] = lh, lv
        for shift in (True, False):
            out = np.zeros(img.shape, np.float32)
            oracle.fir2(lh, lv, img, out, shift)
            fir[f"p{k}_shift{int(shift)}"] = out
    np.savez_compressed(os.path.join(HERE, "fir.npz"), **fir)

    sub = {"image": synth.edge_image(15, 37, 51), "image_u8": synth.u8_image(16, 37, 51)}
    for op in oracle.SUBSAMPLE_OPS:
        out = np.zeros((18, 25), np.float32)
        oracle.subsample(sub["image"], out, op)
        sub[op] = out
    out = np.zeros((18, 25), np.uint8)
    sub["identity1x1_u8"] = oracle.subsample(sub["image_u8"], out, "identity1x1")
    np.savez_compressed(os.path.join(HERE, "subsample.npz"), **sub)


if __name__ == "__main__":
    main()
