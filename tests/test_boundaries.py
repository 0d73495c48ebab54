"""Host-side collision-quad helpers (pbf_amd/boundaries.py): the PNG mask
reader against images encoded here with every scanline filter, and the
reference's force-plate table."""
import struct
import zlib

import numpy as np
import pytest

from pbf_amd import boundaries


def _chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _encode(px: np.ndarray, ctype: int, filters, palette=None) -> bytes:
    """PNG encoder: row y uses filter filters[y % len(filters)]."""
    h, w = px.shape[:2]
    bpp = 1 if px.ndim == 2 else px.shape[2]
    rows = px.reshape(h, w * bpp).astype(np.int64)
    raw = bytearray()
    prev = np.zeros(w * bpp, np.int64)
    for y in range(h):
        ft = filters[y % len(filters)]
        cur = rows[y]
        out = np.zeros_like(cur)
        for x in range(w * bpp):
            a = cur[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = (0, a, b, (a + b) >> 1, _paeth(a, b, c))[ft]
            out[x] = (cur[x] - pred) & 255
        raw.append(ft)
        raw += bytes(out.astype(np.uint8))
        prev = cur
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        png += _chunk(b"PLTE", bytes(palette.astype(np.uint8).ravel()))
    # split the image data over two IDAT chunks
    z = zlib.compress(bytes(raw))
    return png + _chunk(b"IDAT", z[:7]) + _chunk(b"IDAT", z[7:]) + _chunk(b"IEND", b"")


@pytest.mark.parametrize("ctype,channels", [(0, 1), (2, 3), (4, 2), (6, 4)])
def test_png_red_channel_every_filter(tmp_path, ctype, channels):
    rng = np.random.default_rng(ctype)
    px = rng.integers(0, 256, (11, 13, channels)).astype(np.uint8)
    px[3, :, 0] = 0          # a row of zeros and a saturated column
    px[:, 5, 0] = 255
    path = tmp_path / "m.png"
    path.write_bytes(_encode(px if channels > 1 else px[:, :, 0], ctype, [0, 1, 2, 3, 4]))
    got = boundaries.load_mask_png(path)
    assert got.dtype == np.uint32 and got.shape == (11, 13)
    assert np.array_equal(got, px[:, :, 0].astype(np.uint32))


def test_png_palette(tmp_path):
    rng = np.random.default_rng(7)
    pal = rng.integers(0, 256, (6, 3))
    idx = rng.integers(0, 6, (9, 10)).astype(np.uint8)
    path = tmp_path / "p.png"
    path.write_bytes(_encode(idx, 3, [4, 2], palette=pal))
    assert np.array_equal(boundaries.load_mask_png(path), pal[idx, 0].astype(np.uint32))


def test_png_rejects_what_it_does_not_read(tmp_path):
    path = tmp_path / "x.png"
    path.write_bytes(b"not a png")
    with pytest.raises(ValueError):
        boundaries.load_mask_png(path)
    png = _encode(np.zeros((2, 2), np.uint8), 0, [0]).replace(
        struct.pack(">IIBBBBB", 2, 2, 8, 0, 0, 0, 0), struct.pack(">IIBBBBB", 2, 2, 16, 0, 0, 0, 0))
    path.write_bytes(png)
    with pytest.raises(ValueError):
        boundaries.load_mask_png(path)


def test_force_plate_bands_tile_the_mask():
    # each plate's band of mask rows starts where the previous one ends, and
    # the last ends at 9262, the reference mask's height
    plates = boundaries.FORCE_PLATES
    assert len(plates) == 14
    end = 0
    for corner, e0, e1, y, hgt in plates:
        assert y == end and hgt > 0
        end = y + hgt
        n = np.cross(e0, e1)
        assert np.linalg.norm(n) > 0
    assert end == 9262


def test_quad_array_limits():
    arr, k = boundaries.quad_array(boundaries.FORCE_PLATES)
    assert k == 14 and arr[3].mask_y == 1792 and abs(arr[4].e0[0] + 0.0009671) < 1e-9
    _, k = boundaries.quad_array([])
    assert k == 0
    with pytest.raises(ValueError):
        boundaries.quad_array([boundaries.FORCE_PLATES[0]] * 27)
