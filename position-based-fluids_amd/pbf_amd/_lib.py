"""ctypes binding of libpbf_amd.so (the C ABI in include/pbf/pbf.h).

The library is loaded from position-based-fluids_amd/lib/; if it is missing
this module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_ROOT / "lib" / "libpbf_amd.so"

PBF_OK = 0
PBF_E_INVALID = -1
PBF_E_HIP = -2
PBF_E_NOMEM = -3
PBF_E_STATE = -4

PBF_FLAG_SORT_TIMEOUT = 1
PBF_FLAG_FRIENDS_OVERFLOW = 2
PBF_FLAG_NAN = 4
PBF_FLAG_INTERNAL = 8


class PbfParams(ctypes.Structure):
    """Mirror of `struct Parameters` (src/Parameters.hpp:5-56), 124 bytes."""

    _fields_ = [
        ("resetSimOnChange", ctypes.c_int32),
        ("particleCount", ctypes.c_uint32),
        ("xMin", ctypes.c_float),
        ("xMax", ctypes.c_float),
        ("yMin", ctypes.c_float),
        ("yMax", ctypes.c_float),
        ("zMin", ctypes.c_float),
        ("zMax", ctypes.c_float),
        ("waveGenAmp", ctypes.c_float),
        ("waveGenFreq", ctypes.c_float),
        ("waveGenDuty", ctypes.c_float),
        ("timeStep", ctypes.c_float),
        ("simIterations", ctypes.c_uint32),
        ("subSteps", ctypes.c_uint32),
        ("h", ctypes.c_float),
        ("restDensity", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("garvity", ctypes.c_float),
        ("vorticityFactor", ctypes.c_float),
        ("viscosityFactor", ctypes.c_float),
        ("surfaceTenstionK", ctypes.c_float),
        ("surfaceTenstionDist", ctypes.c_float),
        ("friendsCircles", ctypes.c_uint32),
        ("particlesPerCircle", ctypes.c_uint32),
        ("gridBufSize", ctypes.c_uint32),
        ("setupSpacing", ctypes.c_float),
        ("segmentSize", ctypes.c_uint32),
        ("sortIterations", ctypes.c_uint32),
        ("particleRenderSize", ctypes.c_float),
        ("h_2", ctypes.c_float),
        ("EnableCachedBuffers", ctypes.c_bool),
    ]

    def copy(self) -> "PbfParams":
        out = PbfParams()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(PbfParams))
        return out

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class PbfOptions(ctypes.Structure):
    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("use_gravity_vector", ctypes.c_uint32),
        ("walls", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class PbfQuad(ctypes.Structure):
    """pbf_quad: corner, edges E0, E1 and the quad's band of mask rows."""
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("e0", ctypes.c_float * 3),
        ("e1", ctypes.c_float * 3),
        ("mask_y", ctypes.c_int32),
        ("mask_h", ctypes.c_int32),
    ]


PBF_MAX_QUADS = 26


class PbfLogMsg(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint32), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32), ("c", ctypes.c_uint32)]


PBF_LOG_FRIENDS_OVERFLOW = 1
PBF_LOG_NAN = 2
PBF_LOG_SORT_TIMEOUT = 3
PBF_LOG_INTERNAL = 4
PBF_LOG_CAPACITY = 1024
LOG_NAMES = {1: "friends_overflow", 2: "nan", 3: "sort_timeout", 4: "internal"}


class PbfTiming(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("last_ms", ctypes.c_double),
        ("ema_ms", ctypes.c_double),
        ("sum_ms", ctypes.c_double),
        ("calls", ctypes.c_uint64),
    ]


class PbfDevBufs(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p),
        ("predicted", ctypes.c_void_p),
        ("velocities", ctypes.c_void_p),
        ("density", ctypes.c_void_p),
        ("lambda_", ctypes.c_void_p),
        ("omega", ctypes.c_void_p),
        ("stream", ctypes.c_void_p),
        ("n", ctypes.c_uint32),
    ]


class PbfStats(ctypes.Structure):
    _fields_ = [
        ("particles", ctypes.c_uint32),
        ("scenes", ctypes.c_uint32),
        ("total_cells", ctypes.c_uint32),
        ("sort_passes", ctypes.c_uint32),
        ("sort_tiles", ctypes.c_uint32),
        ("unstaged_tiles", ctypes.c_uint32),
        ("workgroup_size", ctypes.c_uint32 * 5),
        ("friends_staged", ctypes.c_uint32),
        ("tiles_per_workgroup", ctypes.c_uint32 * 5),
        ("friends_lanes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 1),
    ]


FAMILIES = ("friends", "lambda", "delta", "viscosity", "vorticity")


class PbfError(RuntimeError):
    """A PBF_E_* return code; the message is pbf_last_error()."""

    def __init__(self, code: int, message: str):
        super().__init__(f"pbf error {code}: {message}")
        self.code = code


class LibraryMissing(ImportError):
    pass


_LIB = None

_VP = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_UP = ctypes.POINTER(ctypes.c_uint32)

_SIGNATURES = {
    "pbf_version": (ctypes.c_char_p, []),
    "pbf_default_options": (None, [ctypes.POINTER(PbfOptions)]),
    "pbf_params_finalize": (None, [ctypes.POINTER(PbfParams)]),
    "pbf_load_parameters": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(PbfParams)]),
    "pbf_create": (ctypes.c_int, [ctypes.POINTER(PbfParams), ctypes.POINTER(PbfOptions), ctypes.POINTER(_VP)]),
    "pbf_create_batch": (ctypes.c_int, [ctypes.POINTER(PbfParams), ctypes.POINTER(PbfOptions), ctypes.c_uint32,
                                        ctypes.POINTER(_VP)]),
    "pbf_scene_count": (ctypes.c_int, [_VP]),
    "pbf_scene_offsets": (ctypes.c_int, [_VP, _UP]),
    "pbf_set_graph": (ctypes.c_int, [_VP, ctypes.c_int]),
    "pbf_reset_timings": (ctypes.c_int, [_VP]),
    "pbf_set_workgroup_sizes": (ctypes.c_int, [_VP, _UP]),
    "pbf_set_friends_mode": (ctypes.c_int, [_VP, ctypes.c_int]),
    "pbf_set_friends_lanes": (ctypes.c_int, [_VP, ctypes.c_uint32]),
    "pbf_set_tiles_per_workgroup": (ctypes.c_int, [_VP, _UP]),
    "pbf_autotune": (ctypes.c_int, [_VP, ctypes.c_uint32, _UP]),
    "pbf_stats": (ctypes.c_int, [_VP, ctypes.POINTER(PbfStats)]),
    "pbf_read_log": (ctypes.c_int, [_VP, ctypes.POINTER(PbfLogMsg), ctypes.c_uint32, _UP,
                                    ctypes.POINTER(ctypes.c_uint64)]),
    "pbf_friend_circle_thresholds": (ctypes.c_int, [ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "pbf_copy_probe": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "pbf_check_fast_math": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), _UP, _UP]),
    "pbf_set_state": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32]),
    "pbf_update_params": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(PbfParams)]),
    "pbf_set_collision_quads": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(PbfQuad), ctypes.c_uint32, _UP,
                                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]),
    "pbf_step": (ctypes.c_int, [_VP, ctypes.c_uint32]),
    "pbf_run": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32]),
    "pbf_run_noalloc": (ctypes.c_int, [ctypes.POINTER(PbfParams), ctypes.POINTER(PbfOptions), _VP, _VP,
                                       ctypes.c_uint32, ctypes.c_uint32, _VP, _VP, _VP]),
    "pbf_sync": (ctypes.c_int, [_VP]),
    "pbf_get_state": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "pbf_get_structure": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "pbf_device_buffers": (ctypes.c_int, [_VP, ctypes.POINTER(PbfDevBufs)]),
    "pbf_set_profiling": (ctypes.c_int, [_VP, ctypes.c_int]),
    "pbf_timings": (ctypes.c_int, [_VP, ctypes.POINTER(PbfTiming), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "pbf_error_flags": (ctypes.c_int, [_VP, _UP, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]),
    "pbf_destroy": (None, [_VP]),
    "pbf_last_error": (ctypes.c_char_p, [_VP]),
    # include/pbf/pbf_adjoint.h
    "pbf_set_scene_coefficients": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _FP]),
    "pbf_get_scene_coefficients": (ctypes.c_int, [_VP, ctypes.c_uint32, _FP]),
    "pbf_adjoint_tape_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "pbf_adjoint_plan": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint64, _UP]),
    "pbf_adjoint_create": (ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_VP)]),
    "pbf_adjoint_forward": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "pbf_adjoint_backward": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "pbf_adjoint_times": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]),
    "pbf_adjoint_overflow": (ctypes.c_int, [_VP, _UP, _UP]),
    "pbf_adjoint_scaled_gradients": (ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.POINTER(ctypes.c_int32)]),
    "pbf_adjoint_destroy": (None, [_VP]),
}


def lib() -> ctypes.CDLL:
    """Loads libpbf_amd.so once; raises LibraryMissing if it was not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        # One HIP runtime per process: PyTorch-ROCm bundles its own
        # libamdhip64.so.7.  Loaded first, it is the one this library binds
        # to (same soname); loaded after this library it would be a second
        # runtime, so import it first whenever it is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def copy_probe(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Sustained plain-copy bandwidth in GB/s (read + write bytes)."""
    g = ctypes.c_double()
    check(lib().pbf_copy_probe(nbytes, reps, ctypes.byref(g)))
    return g.value


def check(rc: int, ctx=None) -> None:
    if rc != PBF_OK:
        msg = lib().pbf_last_error(ctx)
        raise PbfError(rc, msg.decode() if msg else "")


assert ctypes.sizeof(PbfParams) == 124, ctypes.sizeof(PbfParams)
