"""Buffer loads of the fused pyramid step stay inside the input image.

A raw buffer load returns 0 without touching memory when its vector offset
is at or past the buffer's size; the scalar offset is added after that
check, so an address built from a large scalar offset (a row) and a column
past the row's end can reach past the image.  This replays the pyramid
kernel's address arithmetic (tucuda_amd/csrc/subsample.h, pyramid_kernel:
tile origin, column per thread, clamped rows) on the CPU for tightly packed
images and checks that every load that touches memory lies inside the
image's bytes; the form with the row in the scalar offset fails it."""
import numpy as np
import pytest

PYR_TH, PYR_TW = 8, 112


def _pyramid_loads(nrow, ncol, es, r, shift, row_in_soffset=False):
    """(performed, end) of every load: whether it touches memory and the
    byte just past what it reads."""
    R = 4 if r <= 4 else 8 if r <= 8 else 16
    stride_b = ncol * es                                   # tightly packed rows
    in_bytes = (nrow - 1) * stride_b + ncol * es
    onrow, oncol = nrow // 2, ncol // 2
    tiles_y = (onrow + PYR_TH - 1) // PYR_TH
    tiles_x = (oncol + PYR_TW - 1) // PYR_TW
    NW = 2 * (PYR_TH - 1) + 2 * R + 1
    NC = 2 * (PYR_TW - 1) + 2 * R + 1
    o = 0 if shift else r
    tile = np.arange(tiles_x * tiles_y)
    y0 = (tile % tiles_y) * PYR_TH
    x0 = (tile // tiles_y) * PYR_TW
    yc0, xc0 = 2 * y0 + o, 2 * x0 + o
    t = np.arange(NC)
    i = np.arange(NW)
    col = xc0[:, None, None] - R + t[None, None, :]
    row = np.clip(yc0[:, None, None] - R + i[None, :, None], 0, nrow - 1)
    if row_in_soffset:
        voff = np.broadcast_to(col * es, (tile.size, NW, NC)).astype(np.int64)
        soff = np.broadcast_to(row * stride_b, (tile.size, NW, NC)).astype(np.int64)
    else:
        voff = (row * stride_b + col * es).astype(np.int64)
        soff = np.zeros_like(voff)
    voff32 = voff & 0xFFFFFFFF                             # the offset register is 32 bits
    performed = voff32 < in_bytes
    end = voff32 + soff + es
    return performed, end, in_bytes


SHAPES = [(1024, 1024), (1080, 1920), (61, 97), (480, 640), (33, 35)]


@pytest.mark.parametrize("nrow,ncol", SHAPES)
@pytest.mark.parametrize("es", [1, 4])
@pytest.mark.parametrize("r", [3, 8, 16])
@pytest.mark.parametrize("shift", [True, False])
def test_pyramid_loads_inside_image(nrow, ncol, es, r, shift):
    performed, end, in_bytes = _pyramid_loads(nrow, ncol, es, r, shift)
    assert (end[performed] <= in_bytes).all()


def test_row_in_scalar_offset_would_read_past_the_end():
    performed, end, in_bytes = _pyramid_loads(1024, 1024, 4, 16, True, row_in_soffset=True)
    assert (end[performed] > in_bytes).any()
