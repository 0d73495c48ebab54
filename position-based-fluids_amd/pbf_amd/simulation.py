"""Host-side mirror of the reference's Simulation class over the C ABI.

Method names follow src/Simulation.hpp:105-150 (InitBuffers, InitCells,
InitKernels, Step) and the public state flags (bPauseSim); the context is
created lazily by the first of InitBuffers/InitCells/InitKernels, which the
library performs as one call (pbf_create).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (FAMILIES, LOG_NAMES, PbfDevBufs, PbfError, PbfLogMsg, PbfOptions, PbfParams, PbfStats, PbfTiming,
                   check, lib)
from .scenes import default_options

_CIRCLES = 5
_PER_CIRCLE = 50


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Simulation:
    def __init__(self, params: PbfParams, options: PbfOptions | None = None):
        self.params = params.copy()
        self.options = options if options is not None else default_options()
        self._ctx = None
        self.bPauseSim = False
        self.bReadFriendsList = False
        self.fWavePos = 0.0  # passed to computeDelta but unused there (compute_delta.cl:50)

    # -- Simulation::InitBuffers / InitCells / InitKernels ------------------
    def _ensure(self):
        if self._ctx is None:
            ctx = ctypes.c_void_p()
            rc = lib().pbf_create(self.params, self.options, ctypes.byref(ctx))
            if rc != 0:
                raise PbfError(rc, lib().pbf_last_error(None).decode())
            self._ctx = ctx

    def InitBuffers(self):
        self._ensure()

    def InitCells(self):
        self._ensure()

    def InitKernels(self) -> bool:
        self._ensure()
        return True

    # -- state ----------------------------------------------------------------
    @property
    def n(self) -> int:
        return int(self.params.particleCount)

    def set_state(self, x: np.ndarray, v: np.ndarray | None = None):
        """Positions and velocities (float (n, 4) each, n = particleCount; v
        None for zero velocities), checked before any device work."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.n, 4):
            raise ValueError(f"x must have shape ({self.n}, 4), got {x.shape}")
        vv = None
        if v is not None:
            vv = np.ascontiguousarray(v, dtype=np.float32)
            if vv.shape != x.shape:
                raise ValueError(f"v must have shape {x.shape} like x, got {vv.shape}")
        self._ensure()
        check(lib().pbf_set_state(self._ctx, _ptr(x), _ptr(vv), x.shape[0]), self._ctx)

    def update_params(self, params: PbfParams, scene: int = 0):
        """pbf_update_params: the scene's parameters for the next step (a
        reloaded scenario); particleCount, gridBufSize and simIterations must
        stay."""
        self._ensure()
        check(lib().pbf_update_params(self._ctx, int(scene), params), self._ctx)
        if scene == 0 and not hasattr(self, "params_list"):
            self.params = params.copy()

    def set_collision_quads(self, quads, mask=None, edge_base: float = 0.0, scene: int = 0):
        """pbf_set_collision_quads: the scene's collision quads (PbfQuad or
        (origin, e0, e1, mask_y, mask_h) each), applied after the walls, and
        their mask as a (rows, width) texel array (nonzero: solid; None: one
        solid texel, so every quad is solid everywhere).  No quads removes
        them."""
        from .boundaries import quad_array
        self._ensure()
        arr, k = quad_array(quads)
        mp, w, rows = None, 0, 0
        if k:
            m = np.ones((1, 1), np.uint32) if mask is None else np.ascontiguousarray(mask, np.uint32)
            if m.ndim != 2:
                raise ValueError("mask must be a (rows, width) array")
            rows, w = m.shape
            mp = m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        check(lib().pbf_set_collision_quads(self._ctx, int(scene), arr, k, mp, w, rows, float(edge_base)), self._ctx)

    def Step(self):
        """Simulation::Step (src/Simulation.cpp:554-621), asynchronous."""
        self._ensure()
        check(lib().pbf_step(self._ctx, 1 if self.bPauseSim else 0), self._ctx)

    def run(self, steps: int):
        self._ensure()
        check(lib().pbf_run(self._ctx, steps, 1 if self.bPauseSim else 0), self._ctx)

    def sync(self):
        self._ensure()
        check(lib().pbf_sync(self._ctx), self._ctx)

    def get_state(self) -> dict:
        self._ensure()
        n = self.n
        out = {
            "x": np.empty((n, 4), np.float32),
            "v": np.empty((n, 4), np.float32),
            "xs": np.empty((n, 4), np.float32),
            "rho": np.empty(n, np.float32),
            "lambda": np.empty(n, np.float32),
            "omega": np.empty((n, 4), np.float32),
        }
        check(lib().pbf_get_state(self._ctx, _ptr(out["x"]), _ptr(out["v"]), _ptr(out["xs"]), _ptr(out["rho"]),
                                  _ptr(out["lambda"]), _ptr(out["omega"])), self._ctx)
        return out

    def _total_cells(self) -> int:
        return int(self.params.gridBufSize)

    def get_structure(self, friends: bool = True) -> dict:
        self._ensure()
        n = self.n
        out = {
            "keys": np.empty(n, np.uint32),
            "perm": np.empty(n, np.uint32),
            "cells": np.empty(2 * self._total_cells(), np.uint32),
        }
        fr = np.empty(n * _CIRCLES * (1 + _PER_CIRCLE), np.uint32) if friends else None
        check(lib().pbf_get_structure(self._ctx, _ptr(out["keys"]), _ptr(out["perm"]), _ptr(out["cells"]), _ptr(fr)),
              self._ctx)
        if friends:
            out["friends"] = fr
        return out

    def device_buffers(self) -> PbfDevBufs:
        self._ensure()
        b = PbfDevBufs()
        check(lib().pbf_device_buffers(self._ctx, ctypes.byref(b)), self._ctx)
        return b

    # -- OCLPerfMon-style telemetry -------------------------------------------
    def set_profiling(self, enable: bool):
        self._ensure()
        check(lib().pbf_set_profiling(self._ctx, 1 if enable else 0), self._ctx)

    def timings(self) -> dict:
        self._ensure()
        cnt = ctypes.c_int()
        check(lib().pbf_timings(self._ctx, None, 0, ctypes.byref(cnt)), self._ctx)
        arr = (PbfTiming * max(cnt.value, 1))()
        check(lib().pbf_timings(self._ctx, arr, cnt.value, ctypes.byref(cnt)), self._ctx)
        return {
            t.name.decode(): {"last_ms": t.last_ms, "ema_ms": t.ema_ms, "sum_ms": t.sum_ms, "calls": int(t.calls)}
            for t in arr[: cnt.value]
        }

    def reset_timings(self):
        self._ensure()
        check(lib().pbf_reset_timings(self._ctx), self._ctx)

    def set_graph(self, enable: bool):
        """Replay substeps as captured HIP graphs."""
        self._ensure()
        check(lib().pbf_set_graph(self._ctx, 1 if enable else 0), self._ctx)

    def set_workgroup_sizes(self, sizes: dict):
        self._ensure()
        arr = (ctypes.c_uint32 * len(FAMILIES))(*[int(sizes[f]) for f in FAMILIES])
        check(lib().pbf_set_workgroup_sizes(self._ctx, arr), self._ctx)

    FRIENDS_MODES = {"direct": 0, "staged": 1}

    def set_friends_mode(self, mode):
        """Friends build variant (pbf_set_friends_mode): "direct" (the
        default: candidates read from global memory) or "staged" (the tile's
        neighbourhood in LDS); a bool selects staged (True) or direct."""
        self._ensure()
        m = self.FRIENDS_MODES[mode] if isinstance(mode, str) else (1 if mode else 0)
        check(lib().pbf_set_friends_mode(self._ctx, m), self._ctx)

    def set_friends_lanes(self, lanes: int):
        """Lanes per particle of the global-read friends build
        (pbf_set_friends_lanes): 1, 2, 4, 8, 16, or 0 to choose from the
        particle count.  The lists do not depend on it."""
        self._ensure()
        check(lib().pbf_set_friends_lanes(self._ctx, int(lanes)), self._ctx)

    def set_tiles_per_workgroup(self, tiles: dict):
        """Tiles each workgroup of a family handles in turn: 1 (one per tile),
        k > 1, or 0 for a persistent grid of resident workgroups."""
        self._ensure()
        arr = (ctypes.c_uint32 * len(FAMILIES))(*[int(tiles[f]) for f in FAMILIES])
        check(lib().pbf_set_tiles_per_workgroup(self._ctx, arr), self._ctx)

    def autotune(self, steps: int = 3) -> dict:
        """Times every workgroup size per kernel family, then 4 tiles per
        workgroup and a persistent grid; keeps the fastest.  Returns the
        sizes (stats() reports the tiles per workgroup)."""
        self._ensure()
        arr = (ctypes.c_uint32 * len(FAMILIES))()
        check(lib().pbf_autotune(self._ctx, steps, arr), self._ctx)
        return dict(zip(FAMILIES, (int(a) for a in arr)))

    def stats(self) -> dict:
        self._ensure()
        st = PbfStats()
        check(lib().pbf_stats(self._ctx, ctypes.byref(st)), self._ctx)
        out = {name: getattr(st, name) for name, _ in PbfStats._fields_
               if name not in ("workgroup_size", "tiles_per_workgroup", "reserved")}
        out["workgroup_size"] = dict(zip(FAMILIES, (int(a) for a in st.workgroup_size)))
        out["tiles_per_workgroup"] = dict(zip(FAMILIES, (int(a) for a in st.tiles_per_workgroup)))
        return out

    def error_flags(self, reset: bool = False) -> tuple[int, int]:
        self._ensure()
        f = ctypes.c_uint32()
        d = ctypes.c_uint64()
        check(lib().pbf_error_flags(self._ctx, ctypes.byref(f), ctypes.byref(d), 1 if reset else 0), self._ctx)
        return int(f.value), int(d.value)

    def read_log(self, cap: int = 1024) -> tuple[list[dict], int]:
        """Device event log since the previous read (pbf_read_log): messages
        as dicts {event, a, b, c}, oldest first, and how many were lost to
        the ring's wrap."""
        self._ensure()
        arr = (PbfLogMsg * max(cap, 1))()
        cnt, lost = ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().pbf_read_log(self._ctx, arr, cap, ctypes.byref(cnt), ctypes.byref(lost)), self._ctx)
        return ([{"event": LOG_NAMES.get(m.code, str(m.code)), "code": int(m.code), "a": int(m.a), "b": int(m.b),
                  "c": int(m.c)} for m in arr[:cnt.value]], int(lost.value))

    def close(self):
        if self._ctx is not None:
            lib().pbf_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        self._ensure()
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSimulation(Simulation):
    """Independent scenes stepped together (pbf_create_batch): each stage of
    a substep is one launch over every scene.  Arrays are the scenes'
    particles concatenated in scene order; `offsets` gives the ranges."""

    def __init__(self, params_list, options_list=None):
        self.params_list = [p.copy() for p in params_list]
        self.options_list = list(options_list) if options_list is not None else [default_options()
                                                                                   for _ in params_list]
        self.params = self.params_list[0].copy()
        self.params.particleCount = sum(int(p.particleCount) for p in self.params_list)
        self.options = self.options_list[0]
        self._ctx = None
        self.bPauseSim = False
        self.bReadFriendsList = False
        self.fWavePos = 0.0

    def _ensure(self):
        if self._ctx is None:
            S = len(self.params_list)
            P = (PbfParams * S)(*self.params_list)
            O = (PbfOptions * S)(*self.options_list)
            ctx = ctypes.c_void_p()
            rc = lib().pbf_create_batch(P, O, S, ctypes.byref(ctx))
            if rc != 0:
                raise PbfError(rc, lib().pbf_last_error(None).decode())
            self._ctx = ctx

    def _total_cells(self) -> int:
        return sum(int(p.gridBufSize) for p in self.params_list)

    @property
    def offsets(self) -> np.ndarray:
        self._ensure()
        off = np.zeros(len(self.params_list) + 1, np.uint32)
        check(lib().pbf_scene_offsets(self._ctx, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))), self._ctx)
        return off


def load_parameters(text: str) -> tuple[PbfParams, int]:
    """LoadParameters (src/ParamUtils.cpp:24-100) through the C ABI.

    Returns the parameters and the number of unknown keys."""
    p = PbfParams()
    rc = lib().pbf_load_parameters(text.encode(), p)
    if rc < 0:
        raise PbfError(rc, "pbf_load_parameters failed")
    return p, rc
