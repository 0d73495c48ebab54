"""ctypes binding of the CPU oracle (oracle/build/libpbf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "build" / "libpbf_oracle.so"

CIRCLES = 5
PER_CIRCLE = 50


class Cfg(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("h", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("iters", ctypes.c_uint32),
        ("rest_density", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("vorticity", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("st_k", ctypes.c_float),
        ("st_dist", ctypes.c_float),
        ("grid_buf_size", ctypes.c_uint32),
        ("box_lo", ctypes.c_float * 3),
        ("box_hi", ctypes.c_float * 3),
        ("walls", ctypes.c_uint32),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


_L = None


def lib() -> ctypes.CDLL:
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        cp = ctypes.POINTER(Cfg)
        sig = {
            "pbfo_kernel_factors": (None, [ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
            "pbfo_expand_bits": (ctypes.c_uint32, [ctypes.c_int32]),
            "pbfo_cell_hash": (ctypes.c_uint32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32]),
            "pbfo_cell_coord": (ctypes.c_int32, [ctypes.c_float]),
            "pbfo_sinf": (ctypes.c_float, [ctypes.c_float]),
            "pbfo_skip_circle": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_float]),
            "pbfo_predict": (None, [cp, ctypes.c_uint32, vp, vp, vp]),
            "pbfo_keys": (None, [cp, vp, vp]),
            "pbfo_stable_sort": (None, [vp, ctypes.c_uint32, vp, vp]),
            "pbfo_gather": (None, [vp, ctypes.c_uint32, vp, vp]),
            "pbfo_cells": (None, [vp, ctypes.c_uint32, vp]),
            "pbfo_friends": (None, [cp, vp, vp, vp]),
            "pbfo_scaling": (None, [cp, vp, vp, vp, vp]),
            "pbfo_pack": (None, [ctypes.c_uint32, vp, vp, vp]),
            "pbfo_delta": (None, [cp, vp, vp, vp, vp]),
            "pbfo_update_predicted": (None, [ctypes.c_uint32, vp, vp, vp]),
            "pbfo_update_velocities": (None, [cp, vp, vp, vp]),
            "pbfo_viscosity": (None, [cp, vp, vp, vp, vp, vp]),
            "pbfo_vorticity": (None, [cp, vp, vp, vp, vp]),
            "pbfo_create": (vp, [cp]),
            "pbfo_destroy": (None, [vp]),
            "pbfo_set_state": (None, [vp, vp, vp]),
            "pbfo_step": (None, [vp, ctypes.c_uint32]),
            "pbfo_buffer": (vp, [vp, ctypes.c_int]),
            "pbfo_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _L = L
    return _L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cfg_from_params(p, options=None) -> Cfg:
    """Oracle configuration from a PbfParams block (+ PbfOptions)."""
    c = Cfg()
    c.n = p.particleCount
    c.h = p.h
    c.dt = p.timeStep
    c.iters = p.simIterations
    c.rest_density = p.restDensity
    c.epsilon = p.epsilon
    if options is not None and options.use_gravity_vector:
        c.gravity[0], c.gravity[1], c.gravity[2] = options.gravity[0], options.gravity[1], options.gravity[2]
    else:
        c.gravity[0], c.gravity[1], c.gravity[2] = 0.0, -p.garvity, 0.0
    c.vorticity = p.vorticityFactor
    c.viscosity = p.viscosityFactor
    c.st_k = p.surfaceTenstionK
    c.st_dist = p.surfaceTenstionDist
    c.grid_buf_size = p.gridBufSize
    c.box_lo[0], c.box_lo[1], c.box_lo[2] = p.xMin, p.yMin, p.zMin
    c.box_hi[0], c.box_hi[1], c.box_hi[2] = p.xMax, p.yMax, p.zMax
    c.walls = 1 if (options is None or options.walls) else 0
    return c


class OracleSim:
    """Whole-substep oracle (Simulation::Step order)."""

    def __init__(self, cfg: Cfg):
        self.cfg = cfg
        self.n = int(cfg.n)
        self._s = lib().pbfo_create(ctypes.byref(cfg))

    def set_state(self, x, v=None):
        x = np.ascontiguousarray(x, np.float32)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, np.float32)
        lib().pbfo_set_state(self._s, ptr(x), ptr(v))

    def step(self, pause: bool = False):
        lib().pbfo_step(self._s, 1 if pause else 0)

    def _buf(self, which, dtype, count):
        p = lib().pbfo_buffer(self._s, which)
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(
            ctypes.c_float if dtype == np.float32 else ctypes.c_uint32)), shape=(count,))
        return arr.copy()

    def get_state(self) -> dict:
        n = self.n
        return {
            "x": self._buf(0, np.float32, 4 * n).reshape(n, 4),
            "v": self._buf(1, np.float32, 4 * n).reshape(n, 4),
            "xs": self._buf(2, np.float32, 4 * n).reshape(n, 4),
            "rho": self._buf(3, np.float32, n),
            "lambda": self._buf(4, np.float32, n),
            "omega": self._buf(5, np.float32, 4 * n).reshape(n, 4),
        }

    def get_structure(self, friends: bool = True) -> dict:
        n = self.n
        out = {
            "keys": self._buf(6, np.uint32, n),
            "perm": self._buf(7, np.uint32, n),
            "cells": self._buf(8, np.uint32, 2 * int(self.cfg.grid_buf_size)),
        }
        if friends:
            fr = self._buf(9, np.uint32, n * CIRCLES * (1 + PER_CIRCLE))
            counts = fr[: CIRCLES * n].reshape(CIRCLES, n)
            idx = fr[CIRCLES * n:].reshape(CIRCLES, PER_CIRCLE, n)
            mask = np.arange(PER_CIRCLE)[None, :, None] >= counts[:, None, :]
            idx[mask] = 0
            out["friends"] = fr
        return out

    def close(self):
        if self._s:
            lib().pbfo_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_factors(h: float) -> tuple[float, float]:
    a, b = ctypes.c_float(), ctypes.c_float()
    lib().pbfo_kernel_factors(h, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value
