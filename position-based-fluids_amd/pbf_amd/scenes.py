"""Seeded scene generation and the physical -> reference parameter mapping.

The reference kernels are unit-mass: density is POLY6 * sum W with no mass
factor (compute_scaling.cl:99) and XSPH/vorticity have no m/rho_j factor
(apply_viscosity.cl:73-74, apply_vorticity.cl:67-68).  A physical scene with
rest density rho0 and particle mass m = rho0 * (0.5h)^3 therefore maps to

    restDensity     = rho0 / m = (0.5h)^-3
    viscosityFactor = c * m / rho0 = c * (0.5h)^3
    vorticityFactor = eps_vort * m / rho0 = eps_vort * (0.5h)^3

    surfaceTenstionK = -k * h^2

with epsilon, dq, gravity and dt unchanged.  lambda and s_corr are added
together (compute_delta.cl:155) and lambda carries units of length^2
(C / |grad C|^2), so the dimensionless k of the PBF literature is taken in
units of h^2; the literal -0.1 makes s_corr ~10x lambda at h = 0.1 and the
dam break diverges within the first substep.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from ._lib import PbfOptions, PbfParams, lib

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """Counter-based U[0,1) with 24-bit resolution: exact in float32."""
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = splitmix64((counter.astype(np.uint64) + base) & _M64)
    return (z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def dam_break_block(nx: int, ny: int, nz: int, h: float, seed: int, origin=(0.0, 0.0, 0.0),
                    spacing: float = 0.5, jitter: float = 0.01) -> np.ndarray:
    """float32[N,4] positions of an nx*ny*nz block at spacing*h.

    Particle (i, j, k) (x fastest, then y = up, then z) sits at
    origin + (idx + 0.5) * spacing*h + U(-jitter*h, jitter*h)^3, w = 0.
    """
    n = nx * ny * nz
    lin = np.arange(n, dtype=np.int64)
    i = lin % nx
    j = (lin // nx) % ny
    k = lin // (nx * ny)
    d = spacing * h
    out = np.zeros((n, 4), dtype=np.float32)
    for axis, idx in enumerate((i, j, k)):
        u = uniform01(seed, 3 * lin + axis)
        out[:, axis] = (origin[axis] + (idx + 0.5) * d + (2.0 * u - 1.0) * jitter * h).astype(np.float32)
    return out


def reference_block(params: PbfParams) -> np.ndarray:
    """float32[N,4] initial positions of the reference's CreateParticles
    (src/Simulation.cpp:71-101), as the C++ drop-in places them: a cube with
    a = ceil(N^(1/3)) particles per side at spacing d = h * setupSpacing,
    filled y first, then x, then z (particle i at lattice point
    (floor(i/a) mod a, i mod a, floor(i/a^2) mod a)), its x and z centred on
    1/2 and its bottom at y = 0.3; w = 0.  Every operation is binary32 in the
    same order, so the block is bitwise the drop-in's."""
    n = int(params.particleCount)
    a = int(math.ceil(math.pow(n, 1.0 / 3.0)))
    d = np.float32(params.h) * np.float32(params.setupSpacing)
    corner_xz = (np.float32(1.0) - np.float32(a) * d) / np.float32(2.0)
    i = np.arange(n, dtype=np.int64)
    lattice = ((i // a) % a, i % a, (i // (a * a)) % a)
    corner = (corner_xz, np.float32(0.3), corner_xz)
    out = np.zeros((n, 4), np.float32)
    for k in range(3):
        out[:, k] = corner[k] + lattice[k].astype(np.float32) * d
    return out


def morton_grid_size(box_extent, h: float) -> int:
    """Smallest power of two above every in-box Morton id (no hash folding)."""
    cells = [max(1, math.ceil(e / h)) for e in box_extent]
    bits = [max(1, (c - 1).bit_length()) for c in cells]
    m = 0
    for axis, b in enumerate(bits):
        for t in range(b):
            m |= 1 << (3 * t + axis)
    return 1 << m.bit_length()


@dataclasses.dataclass
class Scene:
    params: PbfParams
    options: PbfOptions
    x: np.ndarray          # float32[N,4]
    v: np.ndarray          # float32[N,4]
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.params.particleCount)


def physical_params(n: int, h: float, box_lo, box_hi, *, dt: float = 1.0 / 120.0, iters: int = 4,
                    eps: float = 600.0, k: float = 0.1, dq: float = 0.2, xsph_c: float = 0.01,
                    vort_eps: float = 5e-4, g: float = 9.8, grid_buf_size: int | None = None) -> PbfParams:
    p = PbfParams()
    p.resetSimOnChange = 1
    p.particleCount = n
    p.xMin, p.yMin, p.zMin = (float(a) for a in box_lo)
    p.xMax, p.yMax, p.zMax = (float(a) for a in box_hi)
    p.timeStep = dt
    p.simIterations = iters
    p.subSteps = 1
    p.h = h
    half3 = (0.5 * h) ** 3
    p.restDensity = 1.0 / half3
    p.epsilon = eps
    p.garvity = g
    p.vorticityFactor = vort_eps * half3
    p.viscosityFactor = xsph_c * half3
    p.surfaceTenstionK = -k * h * h
    p.surfaceTenstionDist = dq
    p.setupSpacing = 0.5
    p.segmentSize = 16
    p.sortIterations = 6
    p.particleRenderSize = 1.0
    extent = [hi - lo for lo, hi in zip(box_lo, box_hi)]
    p.gridBufSize = grid_buf_size if grid_buf_size is not None else morton_grid_size(extent, h)
    p.EnableCachedBuffers = False
    lib().pbf_params_finalize(p)
    return p


def default_options(walls: bool = True, gravity=None) -> PbfOptions:
    o = PbfOptions()
    lib().pbf_default_options(o)
    o.walls = 1 if walls else 0
    if gravity is not None:
        o.use_gravity_vector = 1
        o.gravity[0], o.gravity[1], o.gravity[2] = (float(a) for a in gravity)
    return o


def dam_break(nx: int, ny: int, nz: int, h: float, box, *, seed: int = 0, **kw) -> Scene:
    """Dam break: block in the corner of the box [0,box]^3 (y up)."""
    x = dam_break_block(nx, ny, nz, h, seed)
    v = np.zeros_like(x)
    p = physical_params(x.shape[0], h, (0.0, 0.0, 0.0), box, **kw)
    return Scene(p, default_options(), x, v, name=f"dam{nx}x{ny}x{nz}_h{h}")


def inverse_problem(nx: int = 64, ny: int = 32, nz: int = 32, h: float = 0.1, sigma: float = 0.5, seed: int = 0,
                    dt: float = 1.0 / 120.0, iters: int = 4) -> Scene:
    """Block of nx*ny*nz particles at 0.5h in the corner of a box twice its
    extent, with initial velocities N(0, sigma^2) per component (numpy PCG64
    seeded with `seed`): the state whose rollout gives the target positions
    of the inverse problem (default: 65,536 particles, h = 0.1)."""
    x = dam_break_block(nx, ny, nz, h, seed=seed)
    box = (nx * h, ny * h, nz * h)   # block spacing 0.5h: twice its extent
    p = physical_params(x.shape[0], h, (0.0, 0.0, 0.0), box, dt=dt, iters=iters,
                        grid_buf_size=morton_grid_size(box, h))
    v = np.zeros_like(x)
    v[:, :3] = np.random.default_rng(seed).normal(0.0, sigma, (x.shape[0], 3)).astype(np.float32)
    return Scene(p, default_options(), x, v, name=f"inverse{nx}x{ny}x{nz}")


def config(name: str, seed: int = 0) -> Scene:
    """Named dam-break workloads (dt = 1/120 s, 4 solver iterations).

    dam16k  16,384 particles: 32x32x16 block, h=0.1, box 3.2x1.6x1.6
    dam1m   1,048,576 particles: 128x128x64 block, h=0.05, box 12.8x6.4x6.4
    dam4k   4,096 particles: 16x16x16 block, h=0.1, box 1.6x1.6x1.6
    """
    if name == "dam16k":
        return dam_break(32, 32, 16, 0.1, (3.2, 1.6, 1.6), seed=seed)
    if name == "dam1m":
        return dam_break(128, 128, 64, 0.05, (12.8, 6.4, 6.4), seed=seed)
    if name == "dam4k":
        return dam_break(16, 16, 16, 0.1, (1.6, 1.6, 1.6), seed=seed)
    raise KeyError(name)
