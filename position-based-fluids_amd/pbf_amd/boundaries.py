"""Collision quads ("force plates") and their surface mask.

The reference moves particles off a fixed list of quads in its delta stage
(BouncePointQuad, compute_delta.cl:1-42, called at :169-182) wherever a mask
image is set (Simulation::LoadForceMasks, src/Simulation.cpp:257-281: the red
channel of Scene_fp_mask.png as a CL_R / CL_UNSIGNED_INT32 image).  Here the
quads and the mask are per-scene inputs (pbf_set_collision_quads):

    quads = [quad((0, 0, 0), (1, 0, 0), (0, 0, 1), mask_y=0, mask_h=64)]
    sim.set_collision_quads(quads, mask, edge_base=0.0)

FORCE_PLATES is the reference's own quad table and load_mask_png reads a PNG
mask the way LoadForceMasks does, for callers that bring the reference's scene.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

from ._lib import PBF_MAX_QUADS, PbfQuad


def quad(origin, e0, e1, mask_y: int = 0, mask_h: int = 1) -> PbfQuad:
    """One collision quad: corner, edge vectors E0, E1 (the plane normal is
    cross(E0, E1)) and the band of mask rows [mask_y, mask_y + mask_h) that
    quad parameter s in [0, 1] along E0 spans."""
    q = PbfQuad()
    for k in range(3):
        q.origin[k] = float(origin[k])
        q.e0[k] = float(e0[k])
        q.e1[k] = float(e1[k])
    q.mask_y = int(mask_y)
    q.mask_h = int(mask_h)
    return q


def quad_array(quads) -> "ctypes.Array":
    """A ctypes array of PbfQuad from PbfQuad objects or
    (origin, e0, e1, mask_y, mask_h) tuples."""
    quads = [q if isinstance(q, PbfQuad) else quad(*q) for q in quads]
    if len(quads) > PBF_MAX_QUADS:
        raise ValueError(f"{len(quads)} quads > {PBF_MAX_QUADS}")
    return (PbfQuad * max(1, len(quads)))(*quads), len(quads)


# The reference scene's force plates as (corner, E0, E1, mask_y, mask_h), in
# its call order (compute_delta.cl:169-182); they sample a 512-texel-wide mask.
FORCE_PLATES = (
    ((-165.8, -106.4, 28.32), (-15.39, 41.82, 0.0), (52.28, 19.23, 0.0), 0, 409),
    ((-158.5, -126.3, -37.11), (-7.638, 20.76, 66.1), (52.05, 19.15, 0.0), 409, 643),
    ((-21.79, -23.62, -34.55), (89.42, 39.44, 0.0), (-27.26, 61.81, 0.0), 1052, 740),
    ((-39.96, -11.55, 29.95), (-8.877, 49.84, 0.0), (103.3, 18.4, 0.0), 1792, 247),
    ((-180.5, -64.4, -37.25), (-0.0009671, -0.0003558, 66.38), (22.87, -62.17, 0.0), 2039, 512),
    ((35.5, 2.278, -35.25), (0.0, 0.0, 65.91), (21.74, 36.07, 0.0), 2551, 801),
    ((-31.54, 1.768, -8.499), (0.0, 0.0, 15.86), (-15.69, -32.17, 0.0), 3352, 226),
    ((-67.22, -41.65, 7.364), (11.45, 0.0, 0.0), (0.0, 0.0, -15.86), 3578, 369),
    ((-65.91, -42.02, 7.364), (0.0, 0.0, -15.87), (-7.275, 4.603, 0.0), 3947, 943),
    ((-9.05, -4.783, -8.792), (-47.98, -37.3, 0.0), (0.0, 0.0, 16.45), 4890, 1891),
    ((-53.98, -25.37, 6.656), (27.86, 34.59, 0.0), (17.42, -14.03, 0.0), 6781, 1016),
    ((-48.37, -30.96, -7.792), (21.75, -10.61, 0.0), (17.92, 36.76, 0.0), 7797, 303),
    ((-48.28, 38.39, -35.25), (0.0, 0.0, 65.91), (20.93, -47.46, 0.0), 8100, 650),
    ((36.56, 3.007, 30.66), (0.0, 0.0, -65.91), (-64.89, -11.56, 0.0), 8750, 512),
)
FORCE_PLATES_EDGE_BASE = 1.0   # the "1.0 +" of compute_delta.cl:166


def _unfilter(raw: bytes, w: int, h: int, bpp: int) -> np.ndarray:
    """PNG scanline filters 0-4 (PNG specification, section 9).  None and Up
    are vectorised, Sub is a running sum per channel; Average and Paeth walk
    the row."""
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    pos = 0
    for y in range(h):
        ft = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int64)
        pos += 1 + stride
        if ft == 0:
            cur = line
        elif ft == 1:
            cur = np.cumsum(line.reshape(w, bpp), axis=0).reshape(stride) & 255
        elif ft == 2:
            cur = (line + prev) & 255
        elif ft in (3, 4):
            ln, up, cur_l = line.tolist(), prev.tolist(), [0] * stride
            for x in range(stride):
                a = cur_l[x - bpp] if x >= bpp else 0
                b = up[x]
                if ft == 3:
                    pred = (a + b) >> 1
                else:
                    c = up[x - bpp] if x >= bpp else 0
                    p = a + b - c
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                    pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur_l[x] = (ln[x] + pred) & 255
            cur = np.array(cur_l, np.int64)
        else:
            raise ValueError(f"PNG: bad filter type {ft} on row {y}")
        out[y] = cur
        prev = cur
    return out


def load_mask_png(path) -> np.ndarray:
    """The red channel of an 8-bit, non-interlaced PNG as a (rows, width)
    uint32 texel array: what LoadForceMasks uploads (the image is expanded
    to RGBA, green/blue/alpha cleared, and read as one 32-bit texel, so a
    texel equals its red byte; src/Simulation.cpp:257-281)."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, palette, hdr = 8, [], None, None
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            palette = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if hdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, ctype, _, _, interlace = hdr
    channels = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or channels is None or interlace != 0:
        raise ValueError(f"{path}: only 8-bit non-interlaced PNGs are read (depth {depth}, type {ctype}, "
                         f"interlace {interlace})")
    px = _unfilter(zlib.decompress(b"".join(idat)), w, h, channels).reshape(h, w, channels)
    if ctype == 3:
        if palette is None:
            raise ValueError(f"{path}: palette image without PLTE")
        red = palette[px[:, :, 0], 0]
    else:
        red = px[:, :, 0]
    return red.astype(np.uint32)
