"""Seeded ensembles of independent dam-break scenes for the batched path."""
from __future__ import annotations

import math

import numpy as np

from .scenes import Scene, dam_break_block, default_options, physical_params


def ensemble(scenes: int = 64, seed: int = 0, n_min: int = 4096, n_max: int = 131072, h: float = 0.1,
             aspect=(1.0, 4.0), xsph_c=(0.005, 0.05), vort_eps=(0.0, 1e-3), iters: int = 4,
             dt: float = 1.0 / 120.0) -> list[Scene]:
    """Dam-break scenes with log-uniform particle counts.

    Scene s draws (from numpy's PCG64 seeded with `seed`) a target count
    log-uniform in [n_min, n_max], a block aspect (x length over the y and z
    sides) uniform in `aspect`, an XSPH c and a vorticity epsilon uniform in
    their ranges.  The block is nx*ny*nz at 0.5h spacing in the corner of a
    box twice its extent on every axis; its jitter uses seed*1000 + s.
    """
    rng = np.random.default_rng(seed)
    out = []
    for s in range(scenes):
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        a = rng.uniform(*aspect)
        c = rng.uniform(*xsph_c)
        eps = rng.uniform(*vort_eps)
        side = max(2, int(round((target / a) ** (1.0 / 3.0))))
        nx = max(2, int(round(target / (side * side))))
        x = dam_break_block(nx, side, side, h, seed * 1000 + s)
        box = (nx * h, side * h, side * h)   # twice the block extent (block spacing 0.5h)
        p = physical_params(x.shape[0], h, (0.0, 0.0, 0.0), box, dt=dt, iters=iters, xsph_c=c, vort_eps=eps)
        out.append(Scene(p, default_options(), x, np.zeros_like(x), name=f"scene{s}_{nx}x{side}x{side}"))
    return out


def gravity_ensemble(scenes: int = 32, seed: int = 0, n_min: int = 8192, n_max: int = 65536, h: float = 0.1,
                     g: float = 9.8, iters: int = 4, dt: float = 1.0 / 120.0) -> tuple[list[Scene], np.ndarray]:
    """Scenes for batched differentiation: counts log-uniform in [n_min,
    n_max], a cubic-ish block in the corner of a box twice its extent, and a
    per-scene gravity vector of magnitude g with its direction uniform on the
    lower hemisphere (y < 0).  Also returns one target point per scene,
    uniform in the middle half of its box.  Draws from numpy's PCG64 seeded
    with `seed`."""
    from .scenes import morton_grid_size
    rng = np.random.default_rng(seed)
    out, targets = [], []
    for s in range(scenes):
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        side = max(2, int(round(target ** (1.0 / 3.0))))
        nx = max(2, int(round(target / (side * side))))
        x = dam_break_block(nx, side, side, h, seed * 1000 + s)
        box = (nx * h, side * h, side * h)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d[1] = -abs(d[1])
        p = physical_params(x.shape[0], h, (0.0, 0.0, 0.0), box, dt=dt, iters=iters,
                            grid_buf_size=morton_grid_size(box, h))
        o = default_options(gravity=tuple(float(c) for c in g * d))
        out.append(Scene(p, o, x, np.zeros_like(x), name=f"grav{s}_{nx}x{side}x{side}"))
        targets.append([rng.uniform(0.25, 0.75) * e for e in box])
    return out, np.asarray(targets, np.float32)
