"""Collision-quad set-ups shared by the tests: a ramp over the floor of a
scene's box (normal up and towards -x) and a plate facing -x at `plate_x`
of the box width, with a random half-solid mask."""
import numpy as np


def box_quads(params, plate_x=0.55, mask_seed=0):
    lo = np.array([params.xMin, params.yMin, params.zMin], np.float64)
    hi = np.array([params.xMax, params.yMax, params.zMax], np.float64)
    w = hi - lo
    h = float(params.h)
    ramp = ((lo[0], lo[1] + 0.4 * h, lo[2]), (0.0, 0.0, w[2]), (w[0], 0.1 * w[1], 0.0), 0, 24)
    px = lo[0] + plate_x * w[0]
    plate = ((px, lo[1], lo[2]), (0.0, 0.0, w[2]), (0.0, w[1], 0.0), 24, 40)
    mask = (np.random.default_rng(mask_seed).random((64, 32)) < 0.5).astype(np.uint32)
    return [ramp, plate], mask
