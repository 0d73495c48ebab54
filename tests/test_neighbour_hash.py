"""The friends build steps the centre cell's dilated coordinates to its
neighbours' (NeighbourHash, csrc/pbf_device.h) instead of expanding each
neighbour's coordinate (utilities.cl:69-77).  This checks that identity on
every residue the expansion keeps and on the clamped coordinate extremes:
+1 is ((d | ~M) + 1) & M and -1 is (d - 1) & M, M the dilated 10-bit field.
"""
import numpy as np

M = 0x09249249


def expand_bits(x):
    # the reference's spread of the low 10 bits to every third bit, by the
    # definition (bit k of x to bit 3k), not by the shift-and-mask steps
    x = np.asarray(x, dtype=np.int64) & 0x3FF
    out = np.zeros_like(x)
    for k in range(10):
        out |= ((x >> k) & 1) << (3 * k)
    return out.astype(np.uint32)


def test_dilated_steps_match_expansion():
    c = np.concatenate([np.arange(-2048, 2049), [-(1 << 24), (1 << 24), -(1 << 24) + 1, (1 << 24) - 1]])
    e = expand_bits(c).astype(np.uint64)
    up = ((e | (~np.uint64(M) & np.uint64(0xFFFFFFFF))) + 1) & np.uint64(M)
    down = (e - 1) & np.uint64(M)   # e = 0 wraps to 2^64 - 1, masked to M
    assert np.array_equal(up.astype(np.uint32), expand_bits(c + 1))
    assert np.array_equal(down.astype(np.uint32), expand_bits(c - 1))
