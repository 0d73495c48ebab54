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
        ("quads", ctypes.c_void_p),
        ("quad_count", ctypes.c_uint32),
        ("mask", ctypes.c_void_p),
        ("mask_w", ctypes.c_uint32),
        ("mask_rows", ctypes.c_uint32),
        ("edge_base", ctypes.c_float),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


_L = None


def _cpu_has_fma() -> bool:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    return " fma " in line + " "
    except OSError:
        pass
    return True   # no cpuinfo to read: let the load decide


def lib() -> ctypes.CDLL:
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            build()
        if not _cpu_has_fma():
            raise RuntimeError("the oracle is built with -mfma (x86-64 FMA3); this CPU lacks it")
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
            "pbfo_delta_hits": (None, [cp, vp, vp, vp, vp, vp]),
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
            "pbfo_record_walls": (None, [vp, vp]),
            "pbfa_rollout_f64": (None, [cp, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp]),
            "pbfa_jvp_f64": (None, [cp, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "pbfa_jvp_f32": (None, [cp, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "pbfa_tangent_fd": (None, [cp, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       ctypes.c_double, vp, vp]),
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


class Quad(ctypes.Structure):
    _fields_ = [
        ("b", ctypes.c_float * 3),
        ("e0", ctypes.c_float * 3),
        ("e1", ctypes.c_float * 3),
        ("mask_y", ctypes.c_int32),
        ("mask_h", ctypes.c_int32),
    ]


def set_quads(cfg: Cfg, quads, mask=None, edge_base: float = 0.0) -> Cfg:
    """Attaches collision quads ((corner, e0, e1, mask_y, mask_h) tuples) and
    a (rows, width) mask (None: one solid texel) to cfg; cfg keeps them
    alive."""
    arr = (Quad * max(1, len(quads)))()
    for k, (b, e0, e1, my, mh) in enumerate(quads):
        for a in range(3):
            arr[k].b[a], arr[k].e0[a], arr[k].e1[a] = b[a], e0[a], e1[a]
        arr[k].mask_y, arr[k].mask_h = int(my), int(mh)
    m = np.ones((1, 1), np.uint32) if mask is None else np.ascontiguousarray(mask, np.uint32)
    cfg.quads = ctypes.cast(arr, ctypes.c_void_p)
    cfg.quad_count = len(quads)
    cfg.mask = m.ctypes.data
    cfg.mask_rows, cfg.mask_w = m.shape
    cfg.edge_base = edge_base
    cfg._keep_quads = (arr, m)
    return cfg


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


def _opt(a, dtype):
    return None if a is None else ptr(np.ascontiguousarray(a, dtype))


class FrozenRollout:
    """A rollout of `steps` substeps with its integer structure frozen.

    The binary32 oracle runs the substeps from (x0, v0) and records, per
    substep, the sort permutation, the friend lists (reference layout) and the
    wall planes that moved each particle in each iteration.  With that
    structure held fixed the rollout is a smooth function of the initial
    state and of coef = (viscosity, vorticity, g0, g1, g2), which
    pbf_adjoint_ref.cpp evaluates in FP64 (central differences), as FP64 dual
    numbers (JVP) and as binary32 values with FP64 tangents (the derivative of
    the exact binary32 trajectory).  Arrays are (n, 3) in particle-identity
    order."""

    def __init__(self, cfg: Cfg, x0, v0, steps: int):
        self.cfg = cfg
        self.n = n = int(cfg.n)
        self.steps = steps
        iters = int(cfg.iters)
        x4 = np.zeros((n, 4), np.float32)
        v4 = np.zeros((n, 4), np.float32)
        x4[:, :3] = np.asarray(x0, np.float32)[:, :3]
        v4[:, :3] = np.asarray(v0, np.float32)[:, :3]
        self.x0 = x4[:, :3].copy()
        self.v0 = v4[:, :3].copy()
        self.coef0 = np.array([cfg.viscosity, cfg.vorticity, cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]],
                              np.float32)
        sim = OracleSim(cfg)
        sim.set_state(x4, v4)
        hits = np.zeros((iters, n), np.uint32)
        lib().pbfo_record_walls(sim._s, ptr(hits))
        self.perm = np.zeros((steps, n), np.uint32)
        self.friends = np.zeros((steps, n * CIRCLES * (1 + PER_CIRCLE)), np.uint32)
        self.hits = np.zeros((steps, iters, n), np.uint32)
        ids = np.arange(n)
        for s in range(steps):
            sim.step()
            st = sim.get_structure()
            self.perm[s] = st["perm"]
            self.friends[s] = st["friends"]
            self.hits[s] = hits
            ids = ids[st["perm"]]
        lib().pbfo_record_walls(sim._s, None)
        state = sim.get_state()
        sim.close()
        # the binary32 oracle's final state in identity order
        self.xT32 = np.zeros((n, 3), np.float32)
        self.vT32 = np.zeros((n, 3), np.float32)
        self.xT32[ids] = state["x"][:, :3]
        self.vT32[ids] = state["v"][:, :3]
        self.ids = ids

    def _structure(self):
        return (ctypes.byref(self.cfg), self.steps, ptr(self.perm), ptr(self.friends), ptr(self.hits))

    def rollout64(self, x0=None, v0=None, coef=None):
        n = self.n
        x0 = np.ascontiguousarray(self.x0 if x0 is None else x0, np.float64)
        v0 = np.ascontiguousarray(self.v0 if v0 is None else v0, np.float64)
        coef = np.ascontiguousarray(self.coef0 if coef is None else coef, np.float64)
        xT = np.zeros((n, 3))
        vT = np.zeros((n, 3))
        lib().pbfa_rollout_f64(*self._structure(), ptr(x0), ptr(v0), ptr(coef), ptr(xT), ptr(vT))
        return xT, vT

    def jvp64(self, dx0=None, dv0=None, dcoef=None, x0=None, v0=None, coef=None):
        n = self.n
        x0 = np.ascontiguousarray(self.x0 if x0 is None else x0, np.float64)
        v0 = np.ascontiguousarray(self.v0 if v0 is None else v0, np.float64)
        coef = np.ascontiguousarray(self.coef0 if coef is None else coef, np.float64)
        out = [np.zeros((n, 3)) for _ in range(4)]
        lib().pbfa_jvp_f64(*self._structure(), ptr(x0), ptr(v0), ptr(coef), _opt(dx0, np.float64),
                           _opt(dv0, np.float64), _opt(dcoef, np.float64), *[ptr(o) for o in out])
        return tuple(out)

    def jvp32(self, dx0=None, dv0=None, dcoef=None):
        n = self.n
        x0 = np.ascontiguousarray(self.x0, np.float32)
        v0 = np.ascontiguousarray(self.v0, np.float32)
        coef = np.ascontiguousarray(self.coef0, np.float32)
        xT = np.zeros((n, 3), np.float32)
        vT = np.zeros((n, 3), np.float32)
        dxT = np.zeros((n, 3))
        dvT = np.zeros((n, 3))
        lib().pbfa_jvp_f32(*self._structure(), ptr(x0), ptr(v0), ptr(coef), _opt(dx0, np.float64),
                           _opt(dv0, np.float64), _opt(dcoef, np.float64), ptr(xT), ptr(vT), ptr(dxT), ptr(dvT))
        return xT, vT, dxT, dvT


    def tangent_fd(self, dx0=None, dv0=None, dcoef=None, delta: float = 1e-6):
        """FP64 central differences of each substep, chained along the
        binary32 trajectory (perturbations of `delta` in the largest
        component)."""
        n = self.n
        x0 = np.ascontiguousarray(self.x0, np.float32)
        v0 = np.ascontiguousarray(self.v0, np.float32)
        coef = np.ascontiguousarray(self.coef0, np.float32)
        dxT = np.zeros((n, 3))
        dvT = np.zeros((n, 3))
        lib().pbfa_tangent_fd(*self._structure(), ptr(x0), ptr(v0), ptr(coef), _opt(dx0, np.float64),
                              _opt(dv0, np.float64), _opt(dcoef, np.float64), delta, ptr(dxT), ptr(dvT))
        return dxT, dvT


def kernel_factors(h: float) -> tuple[float, float]:
    a, b = ctypes.c_float(), ctypes.c_float()
    lib().pbfo_kernel_factors(h, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value
