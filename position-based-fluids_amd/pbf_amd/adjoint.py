"""Differentiable rollouts over the hand-written HIP adjoint.

A Rollout wraps a pbf_adjoint (include/pbf/pbf_adjoint.h) on a Simulation or
BatchSimulation context: `substeps` calls of the substep from an initial
state, differentiable with respect to the initial positions and velocities
and to each scene's coefficients coef = (viscosityFactor, vorticityFactor,
g0, g1, g2) with the sort order, friend lists and wall branches frozen at the
values the forward produced.

    roll = Rollout(sim, substeps=60, interval=10)
    xT, vT = roll(x0, v0, coef)        # torch tensors on the GPU, autograd-enabled
    loss = ((xT - target) ** 2).sum(1).mean()
    loss.backward()                     # x0.grad, v0.grad, coef.grad

The autograd node is the libtorch torch::autograd::Function of
csrc/pbf_torch.cpp (module pbf_amd.pbf_torch).  Physical XSPH c and
vorticity epsilon map to the reference factors by the constant (0.5h)^3
(scenes.physical_params), so gradients with respect to them follow by the
chain rule through that factor (see physical_coef).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def check_float4(t, n: int, what: str) -> None:
    """Raises ValueError unless t is a contiguous float32 (n, 4) GPU tensor:
    the library reads or writes n float4 rows through its data pointer."""
    shape = tuple(getattr(t, "shape", ()))
    if shape != (n, 4):
        raise ValueError(f"{what} must have shape ({n}, 4), got {shape}")
    if str(getattr(t, "dtype", "")) != "torch.float32":
        raise ValueError(f"{what} must be float32, got {getattr(t, 'dtype', type(t))}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{what} must be a GPU tensor")


def tape_bytes(sim) -> tuple[int, int]:
    """Device bytes of one recorded substep and of one checkpoint."""
    sim._ensure()
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().pbf_adjoint_tape_bytes(sim._ctx, ctypes.byref(a), ctypes.byref(b)), sim._ctx)
    return a.value, b.value


def plan_interval(sim, substeps: int, budget_bytes: int = 64 << 30) -> int:
    """Checkpoint interval under a device-memory budget (pbf_adjoint_plan)."""
    sim._ensure()
    out = ctypes.c_uint32()
    check(lib().pbf_adjoint_plan(sim._ctx, substeps, budget_bytes, ctypes.byref(out)), sim._ctx)
    return int(out.value)


def physical_coef(h: float, xsph_c, vort_eps, gravity):
    """coef row (reference units) from physical XSPH c, vorticity epsilon and a
    gravity vector; works on floats or torch tensors (differentiably)."""
    half3 = (0.5 * h) ** 3
    import torch
    parts = [torch.as_tensor(xsph_c, dtype=torch.float32) * half3,
             torch.as_tensor(vort_eps, dtype=torch.float32) * half3]
    g = torch.as_tensor(gravity, dtype=torch.float32)
    return torch.cat([torch.stack(parts), g.reshape(3)])


class Rollout:
    def __init__(self, sim, substeps: int, interval: int | None = None, budget_bytes: int = 64 << 30):
        sim._ensure()
        self.sim = sim
        self._ctx = sim._ctx
        self.substeps = int(substeps)
        self.interval = int(interval) if interval else plan_interval(sim, substeps, budget_bytes)
        a = ctypes.c_void_p()
        check(lib().pbf_adjoint_create(self._ctx, self.substeps, self.interval, ctypes.byref(a)), self._ctx)
        self._a = a
        self.scenes = int(lib().pbf_scene_count(self._ctx))
        self.n = sim.n

    # -- coefficients ---------------------------------------------------------
    def coefficients(self) -> np.ndarray:
        out = np.zeros((self.scenes, 5), np.float32)
        for s in range(self.scenes):
            check(lib().pbf_get_scene_coefficients(self._ctx, s, out[s].ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                  self._ctx)
        return out

    def set_coefficients(self, coef) -> None:
        coef = np.ascontiguousarray(np.asarray(coef, np.float32).reshape(self.scenes, 5))
        for s in range(self.scenes):
            g = coef[s, 2:5].copy()
            check(lib().pbf_set_scene_coefficients(self._ctx, s, float(coef[s, 0]), float(coef[s, 1]),
                                                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), self._ctx)

    # -- autograd -------------------------------------------------------------
    def __call__(self, x0, v0, coef=None):
        """(xT, vT) = rollout(x0, v0, coef) as a torch autograd node."""
        import torch

        from . import pbf_torch
        if coef is None:
            coef = torch.from_numpy(self.coefficients())
        return tuple(pbf_torch.rollout(self._a.value, self._ctx.value, x0, v0, coef))

    # -- raw device calls (torch tensors as device memory) --------------------
    def forward_raw(self, x0, v0):
        """x0, v0: float32 (n, 4) contiguous GPU tensors; returns (xT, vT)."""
        import torch
        check_float4(x0, self.n, "x0")
        check_float4(v0, self.n, "v0")
        xT = torch.empty_like(x0)
        vT = torch.empty_like(v0)
        torch.cuda.synchronize()
        check(lib().pbf_adjoint_forward(self._a, x0.data_ptr(), v0.data_ptr(), xT.data_ptr(), vT.data_ptr()),
              self._ctx)
        return xT, vT

    def backward_raw(self, gxT=None, gvT=None):
        """Returns (dL/dx0, dL/dv0, dL/dcoef) for the last forward_raw."""
        import torch
        if gxT is not None:
            check_float4(gxT, self.n, "gxT")
        if gvT is not None:
            check_float4(gvT, self.n, "gvT")
        dev = torch.device("cuda")
        gx0 = torch.empty((self.n, 4), dtype=torch.float32, device=dev)
        gv0 = torch.empty((self.n, 4), dtype=torch.float32, device=dev)
        gc = torch.empty((self.scenes, 5), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        check(lib().pbf_adjoint_backward(self._a, None if gxT is None else gxT.data_ptr(),
                                         None if gvT is None else gvT.data_ptr(), gx0.data_ptr(), gv0.data_ptr(),
                                         gc.data_ptr()), self._ctx)
        return gx0, gv0, gc

    def overflow(self) -> np.ndarray:
        """Per scene, bits of the gradients of the last backward that exceed
        binary32 (returned as +-inf): 1 coefficient, 2 dL/dx0 or dL/dv0."""
        k = ctypes.c_uint32()
        out = np.zeros(self.scenes, np.uint32)
        check(lib().pbf_adjoint_overflow(self._a, ctypes.byref(k), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              self._ctx)
        return out

    def scaled_gradients(self):
        """The last backward's gradients as (gx0, gv0, gcoef) mantissas (GPU
        tensors) and a per-scene exponent array: true = mantissa * 2^expo."""
        import torch
        dev = torch.device("cuda")
        gx0 = torch.empty((self.n, 4), dtype=torch.float32, device=dev)
        gv0 = torch.empty((self.n, 4), dtype=torch.float32, device=dev)
        gc = torch.empty((self.scenes, 5), dtype=torch.float32, device=dev)
        expo = np.zeros(self.scenes, np.int32)
        torch.cuda.synchronize()
        check(lib().pbf_adjoint_scaled_gradients(self._a, gx0.data_ptr(), gv0.data_ptr(), gc.data_ptr(),
                                                 expo.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), self._ctx)
        return gx0, gv0, gc, expo

    def times(self) -> dict:
        f, b, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().pbf_adjoint_times(self._a, ctypes.byref(f), ctypes.byref(b), ctypes.byref(r)), self._ctx)
        return {"forward_ms": f.value, "backward_ms": b.value, "replay_ms": r.value}

    def close(self):
        if getattr(self, "_a", None) is not None and self._a.value:
            lib().pbf_adjoint_destroy(self._a)
            self._a = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
