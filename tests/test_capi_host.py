"""The C ABI without a GPU: the library loads, exports every entry point
include/tucu.h declares, and its host-only functions (lobe generation, size
rule, argument checks, row stride) behave as the reference's host code."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from tucuda_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "tucu.h")).read()
    return sorted(set(re.findall(r"TUCU_API\s+[\w\s\*]+?\b(tucu_\w+)\s*\(", text)))


def test_header_lists_every_export():
    assert declared_entry_points() == sorted(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    for name in declared_entry_points():
        assert hasattr(lib, name), name
    out = os.popen(f"nm -D --defined-only {_capi.LIB_PATH}").read()
    exported = set(re.findall(r"\bT (tucu_\w+)", out))
    assert exported == set(declared_entry_points())


def test_version_and_errors():
    lib = _capi.load()
    assert lib.tucu_version() >= 100
    assert lib.tucu_fir2_check(17, 17) == 0
    assert lib.tucu_fir2_check(18, 3) == _capi.TUCU_ERR_LOBE_TOO_LARGE
    assert b"too large lobe size" in lib.tucu_last_error()
    with pytest.raises(_capi.LobeTooLarge):
        _capi.check(lib.tucu_fir2_check(3, 20))
    assert lib.tucu_fir2_check(-1, 3) == _capi.TUCU_ERR_INVALID


def test_row_stride_256_byte_rows():
    lib = _capi.load()
    for ncol, es in [(1, 1), (255, 1), (256, 1), (257, 1), (640, 4), (1920, 4), (100, 2), (65, 4)]:
        s = lib.tucu_row_stride(ncol, es)
        assert s >= ncol and (s * es) % 256 == 0 and (s - ncol) * es < 256


def test_host_lobes_equal_oracle_bitwise():
    """tucu_gauss2_lobes (reference rule) reproduces the oracle's float
    lobes bit for bit over a dense sigma sweep."""
    import tucuda_amd
    for sigma in np.linspace(0.5, 5.0, 181):
        s = float(np.float32(sigma))
        ours = tucuda_amd.gauss_lobes(s)
        ref = oracle.gauss_lobes(s)
        for a, b in zip(ours, ref):
            assert a.size == b.size and (a.view(np.uint32) == b.view(np.uint32)).all(), sigma


def test_lobe_size_equals_oracle():
    import tucuda_amd
    rng = np.random.default_rng(0)
    for _ in range(300):
        lobe = (rng.standard_normal(17) * np.exp(-rng.uniform(0, 1) * np.arange(17))).astype(np.float32)
        for even in (True, False):
            assert tucuda_amd.lobe_size(lobe, even) == oracle.lobe_size(lobe, even)


def test_ceil4sigma_rule_equals_oracle_bitwise():
    """Extension rule: radius ceil(4 sigma) capped at 16, the reference's
    normalisations (unit DC gain for lobe 0, quadratic response 2); the host
    library's float lobes equal the oracle's FP32 ceil(4 sigma) lobes bit for
    bit (the oracle's rule is pinned in tests/test_oracle_lobes.py)."""
    import tucuda_amd
    for sigma in np.linspace(0.5, 5.0, 181):
        s = float(np.float32(sigma))
        ours = tucuda_amd.gauss_lobes(s, tucuda_amd.TUCU_RULE_CEIL4SIGMA)
        ref = oracle.gauss_lobes(s, "f32", "ceil4")
        for a, b in zip(ours, ref):
            assert a.size == b.size and (a.view(np.uint32) == b.view(np.uint32)).all(), sigma
    for sigma in (0.8, 1.0, 1.3, 2.2, 3.0, 3.99, 4.0):
        l0, l1, l2 = tucuda_amd.gauss_lobes(sigma, tucuda_amd.TUCU_RULE_CEIL4SIGMA)
        r = min(16, int(np.ceil(4 * np.float32(sigma))))
        assert (l0.size, l1.size, l2.size) == (r + 1, r, r + 1)
        k0 = np.concatenate([l0[:-1], l0[-1:], l0[:-1][::-1]]).astype(np.float64)
        assert abs(k0.sum() - 1) < 2e-6
        dist = np.arange(r, 0, -1, dtype=np.float64)
        assert abs(2 * np.dot(l2[:-1].astype(np.float64), dist ** 2) - 2.0) < 1e-4
        assert (l1 < 0).all()


def _img(data, stride, nrow, ncol, dt):
    return _capi.TucuImage(data, stride, nrow, ncol, dt, 0)


def test_subsample_and_pyramid_argument_checks():
    """Argument checks of tucu_subsample and tucu_gauss2_pyramid, which
    return before any device work: unknown operator, integer input to an
    arithmetic operator, output too small, too-small input (a no-op,
    algorithm.h:227-233), misaligned data, unsupported lobe length."""
    lib = _capi.load()
    fake = 0x10000                                   # never dereferenced on these paths
    f32, u8 = _capi.TUCU_F32, _capi.TUCU_U8
    src = _img(fake, 64, 32, 64, f32)
    dst = _img(fake, 32, 16, 32, f32)
    assert lib.tucu_subsample(ctypes.byref(src), ctypes.byref(dst), 12, None) == _capi.TUCU_ERR_INVALID
    assert b"unknown subsample operator" in lib.tucu_last_error()
    assert lib.tucu_subsample(ctypes.byref(src), ctypes.byref(dst), -1, None) == _capi.TUCU_ERR_INVALID
    src8 = _img(fake, 64, 32, 64, u8)
    assert lib.tucu_subsample(ctypes.byref(src8), ctypes.byref(dst), 1, None) == _capi.TUCU_ERR_INVALID
    assert b"floating-point" in lib.tucu_last_error()
    small = _img(fake, 32, 15, 32, f32)
    assert lib.tucu_subsample(ctypes.byref(src), ctypes.byref(small), 0, None) == _capi.TUCU_ERR_INVALID
    assert b"smaller" in lib.tucu_last_error()
    one_row = _img(fake, 64, 1, 64, f32)
    assert lib.tucu_subsample(ctypes.byref(one_row), ctypes.byref(dst), 11, None) == _capi.TUCU_OK
    odd = _img(fake + 2, 64, 32, 64, f32)
    assert lib.tucu_subsample(ctypes.byref(odd), ctypes.byref(dst), 0, None) == _capi.TUCU_ERR_INVALID
    assert b"aligned" in lib.tucu_last_error()

    lobe = np.array([0.1, 0.2, 0.4], np.float32)
    fp = lobe.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.tucu_gauss2_pyramid(fp, 1, ctypes.byref(src), ctypes.byref(dst), 1, None) == \
        _capi.TUCU_ERR_UNSUPPORTED_LOBE
    assert lib.tucu_gauss2_pyramid(fp, 18, ctypes.byref(src), ctypes.byref(dst), 1, None) == \
        _capi.TUCU_ERR_UNSUPPORTED_LOBE
    assert lib.tucu_gauss2_pyramid(fp, 3, ctypes.byref(src), ctypes.byref(small), 1, None) == _capi.TUCU_ERR_INVALID
    assert lib.tucu_gauss2_pyramid(fp, 3, ctypes.byref(one_row), ctypes.byref(dst), 1, None) == _capi.TUCU_OK
    assert lib.tucu_gauss2_pyramid(None, 3, ctypes.byref(src), ctypes.byref(dst), 1, None) == _capi.TUCU_ERR_INVALID


def test_memcpy2d_argument_checks():
    lib = _capi.load()
    buf = ctypes.create_string_buffer(64)
    assert lib.tucu_memcpy2d(buf, 16, buf, 16, 8, 2, 3) == _capi.TUCU_ERR_INVALID
    assert b"copy kind" in lib.tucu_last_error()
    assert lib.tucu_memcpy2d(buf, 4, buf, 16, 8, 2, 0) == _capi.TUCU_ERR_INVALID
    assert b"pitch" in lib.tucu_last_error()
    assert lib.tucu_memcpy2d(None, 16, buf, 16, 8, 2, 0) == _capi.TUCU_ERR_INVALID
    assert lib.tucu_memcpy2d(buf, 16, buf, 16, 0, 2, 0) == _capi.TUCU_OK      # nothing to copy


def test_lobe_size_null_is_negative():
    lib = _capi.load()
    assert lib.tucu_lobe_size(None, 0) == -_capi.TUCU_ERR_INVALID
    assert b"null lobe" in lib.tucu_last_error()


def test_server_admission_rule():
    """tucu_server_admit_bytes: grow by max(need, step) clipped to the free
    device memory less the reserve and to the pool cap; 0 (wait) when the
    request fits in neither."""
    lib = _capi.load()
    f = lib.tucu_server_admit_bytes
    G, M = 1 << 30, 1 << 20
    assert f(10 * M, 50 * G, G, 0, 0, 256 * M) == 256 * M          # the growth step
    assert f(300 * M, 50 * G, G, 0, 0, 256 * M) == 300 * M         # a request larger than the step
    assert f(10 * M, G + 100 * M, G, 0, 0, 256 * M) == 100 * M     # clipped to free - reserve
    assert f(200 * M, G + 100 * M, G, 0, 0, 256 * M) == 0          # does not fit: wait
    assert f(10 * M, 500 * M, G, 0, 0, 256 * M) == 0               # less free than the reserve
    assert f(10 * M, 50 * G, G, 4 * G, 4 * G + 64 * M, 256 * M) == 64 * M   # clipped to the cap
    assert f(100 * M, 50 * G, G, 4 * G, 4 * G + 64 * M, 256 * M) == 0       # beyond the cap
    assert f(10 * M, 50 * G, G, 4 * G, 4 * G, 256 * M) == 0                 # a fixed pool
    assert f(0, 50 * G, G, 0, 0, 256 * M) == 0


def test_icia_moment_argument_checks():
    """Argument checks of tucu_icia_moment, which return before any device
    work: unknown model, lobe lengths outside 1..16, the workspace size
    query, a too-small workspace, and a workspace that overlaps the image or
    the moment."""
    import tucuda_amd
    lib = _capi.load()
    fp = ctypes.POINTER(ctypes.c_float)
    l0, l1, _ = tucuda_amd.gauss_lobes(2.0)
    p0, p1 = l0.ctypes.data_as(fp), l1.ctypes.data_as(fp)
    img = 0x100000                                   # never dereferenced on these paths
    src = _img(img, 512, 200, 300, _capi.TUCU_F32)
    need = ctypes.c_size_t(0)

    def call(model, moment, ws, nbytes, n0=l0.size, n1=l1.size):
        size = ctypes.c_size_t(nbytes)
        return lib.tucu_icia_moment(p0, n0, p1, n1, ctypes.byref(src), model, ctypes.c_void_p(moment),
                                    ctypes.c_void_p(ws), ctypes.byref(size), None)

    assert lib.tucu_icia_moment(p0, l0.size, p1, l1.size, ctypes.byref(src), 6, None, None,
                                ctypes.byref(need), None) == _capi.TUCU_OK
    assert need.value > 0
    n = need.value
    other = 0x40000000                               # far from the image
    assert call(7, other, other + 4096, n) == _capi.TUCU_ERR_INVALID
    assert call(6, other, other + 4096, n, n0=1) == _capi.TUCU_ERR_UNSUPPORTED_LOBE
    assert call(6, other, other + 4096, n, n1=17) == _capi.TUCU_ERR_UNSUPPORTED_LOBE
    assert call(6, other, other + 4096, n - 1) == _capi.TUCU_ERR_INVALID
    assert call(6, other + 256, other, n) == _capi.TUCU_ERR_INVALID                  # moment inside the workspace
    assert b"moment overlaps the workspace" in lib.tucu_last_error()
    assert call(6, other, img + 199 * 512 * 4, n) == _capi.TUCU_ERR_INVALID          # workspace on the last row
    assert b"workspace overlaps the input" in lib.tucu_last_error()
