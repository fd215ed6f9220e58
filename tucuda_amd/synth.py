"""Seeded synthetic inputs for the benchmark and the tests.

The generator is SplitMix64 (Steele, Lea and Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014): element ``i`` of stream
``seed`` is ``mix(seed + (i + 1) * 0x9E3779B97F4A7C15)``.  Uniform floats
take the top 24 bits (``(z >> 40) * 2**-24``), bytes the top 8 bits, so
the same numbers can be produced from any language.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Elements ``offset .. offset+n-1`` of the SplitMix64 stream ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """float64 values k * 2**-24, k uniform in [0, 2**24)."""
    return (splitmix64(seed, n, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def uniform_image(seed: int, nrow: int, ncol: int, hi: float = 255.0) -> np.ndarray:
    """float32 image, pixels uniform in [0, hi)."""
    return (uniform01(seed, nrow * ncol) * hi).astype(np.float32).reshape(nrow, ncol)


def u8_image(seed: int, nrow: int, ncol: int) -> np.ndarray:
    """uint8 image, pixels uniform in {0..255}."""
    return (splitmix64(seed, nrow * ncol) >> np.uint64(56)).astype(np.uint8).reshape(nrow, ncol)


def edge_image(seed: int, nrow: int, ncol: int, nedges: int = 8, amplitude: float = 64.0) -> np.ndarray:
    """Uniform [0,255) float32 image plus ``nedges`` step edges of height
    ``amplitude``.  Edge k is the half-plane (x-cx)cos(t) + (y-cy)sin(t) > 0
    with t, cx, cy drawn from the same stream after the pixels."""
    img = uniform01(seed, nrow * ncol).reshape(nrow, ncol) * 255.0
    par = uniform01(seed, 3 * nedges, offset=nrow * ncol).reshape(nedges, 3)
    yy, xx = np.mgrid[0:nrow, 0:ncol].astype(np.float64)
    for t, u, v in par:
        th, cx, cy = 2.0 * np.pi * t, u * ncol, v * nrow
        img += amplitude * (((xx - cx) * np.cos(th) + (yy - cy) * np.sin(th)) > 0.0)
    return img.astype(np.float32)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """N(0,1) samples by Box-Muller over pairs of uniforms."""
    u = uniform01(seed, 2 * n, offset).reshape(n, 2)
    u1 = np.maximum(u[:, 0], 1.0 / (1 << 25))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[:, 1])
