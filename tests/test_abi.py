"""C ABI checks that need no GPU: the library loads, exports every declared
symbol, keeps the reference Parameters layout, and parses .par text like
src/ParamUtils.cpp:24-100."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import pbf_amd
from pbf_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for hdr in (ROOT / "include" / "pbf").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = pbf_amd.lib()
    names = declared_functions()
    assert "pbf_step" in names and "pbf_create" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_parameters_layout_matches_reference():
    # src/Parameters.hpp:5-56 on x86-64: 124 bytes
    P = _lib.PbfParams
    assert ctypes.sizeof(P) == 124
    off = {name: getattr(P, name).offset for name, _ in P._fields_}
    assert off["particleCount"] == 4
    assert off["timeStep"] == 44
    assert off["simIterations"] == 48
    assert off["h"] == 56
    assert off["garvity"] == 68
    assert off["surfaceTenstionDist"] == 84
    assert off["gridBufSize"] == 96
    assert off["h_2"] == 116
    assert off["EnableCachedBuffers"] == 120


PAR = """# Simulation settings
ParticleCount           50000
SetupSpacing            0.55
ResetSimOnChange        1
XMin                    -30.0   # trailing comment
xmax                    90.0
TimeStep                0.025
SubSteps                1
SimIterations           3
SmoothLen               2.0
RestDensity             1.0
Garvity                 10.0
Epsilon                 5.0
VorticityFactor         5.0
ViscosityFactor         0.01
SurfaceTenstionK       -0.1
SurfaceTenstionDist     0.1
GridBufferSize          128000
NotAKey                 3
EnableCachedBuffers     1
"""


def test_load_parameters_like_reference():
    p, unknown = pbf_amd.load_parameters(PAR)
    assert unknown == 1
    assert p.particleCount == 50000
    assert p.xMin == -30.0 and p.xMax == 90.0
    assert p.simIterations == 3 and p.h == 2.0
    assert p.gridBufSize == 128000
    assert abs(p.surfaceTenstionK + 0.1) < 1e-7
    assert p.EnableCachedBuffers
    # computed fields (ParamUtils.cpp:97-99)
    assert p.h_2 == 4.0 and p.friendsCircles == 5 and p.particlesPerCircle == 50


def test_load_parameters_drops_unterminated_last_line():
    # ParamUtils.cpp:35-39: getline on a last line without '\n' leaves the
    # stream not good(), which ends the parse before that line is used
    p, _ = pbf_amd.load_parameters("SmoothLen 0.5\nSimIterations 7")
    assert p.h == 0.5 and p.simIterations == 0


def test_create_rejects_bad_parameters_without_gpu():
    # validation happens before any HIP call
    scene = pbf_amd.config("dam4k")
    p = scene.params.copy()
    p.simIterations = 0
    with pytest.raises(pbf_amd.PbfError) as e:
        pbf_amd.Simulation(p).InitBuffers()
    assert e.value.code == _lib.PBF_E_INVALID
    p = scene.params.copy()
    p.friendsCircles = 4
    with pytest.raises(pbf_amd.PbfError):
        pbf_amd.Simulation(p).InitBuffers()


def test_morton_grid_size():
    assert pbf_amd.morton_grid_size((3.2, 1.6, 1.6), 0.1) == 1 << 13
    assert pbf_amd.morton_grid_size((12.8, 6.4, 6.4), 0.05) == 1 << 22


def test_scene_generator_is_deterministic():
    a = pbf_amd.dam_break_block(8, 4, 2, 0.1, seed=3)
    b = pbf_amd.dam_break_block(8, 4, 2, 0.1, seed=3)
    c = pbf_amd.dam_break_block(8, 4, 2, 0.1, seed=4)
    assert (a == b).all() and not (a == c).all()
    d = a[:, :3] - ((pbf_amd.dam_break_block(8, 4, 2, 0.1, seed=3, jitter=0.0))[:, :3])
    assert abs(d).max() <= 0.01 * 0.1 + 1e-7


@pytest.mark.parametrize("h", [0.1, 0.05, 0.125, 2.0, 0.0731])
def test_friend_circle_thresholds_match_formula(h):
    """The kernels' r^2 thresholds reproduce min(int(a^3*5), 4),
    a = max(0, (sqrt(r^2) - 0.3h)/(h - 0.3h)) (build_friends_list.cl:75-76)
    exactly, including the binary32 values next to every threshold."""
    t = (ctypes.c_float * 4)()
    assert pbf_amd.lib().pbf_friend_circle_thresholds(h, t) == 0
    t = np.array(list(t), np.float32)
    hf = np.float32(h)
    h2 = hf * hf
    min_r = np.float32(0.3) * hf
    span = hf - min_r

    def circle(r2):
        a = (np.sqrt(r2) - min_r) / span
        a = np.maximum(np.float32(0.0), a)
        c = (a * a * a * np.float32(5.0)).astype(np.int32)
        return np.minimum(c, 4)

    rng = np.random.default_rng(0)
    r2 = (rng.random(200000) * float(h2)).astype(np.float32)
    near = []
    for v in t[np.isfinite(t)]:
        bits = int(np.array([v], np.float32).view(np.uint32)[0])
        near += [np.array([bits + d], np.uint32).view(np.float32)[0] for d in range(-3, 4)]
    r2 = np.concatenate([r2, np.array(near, np.float32), np.array([0.0], np.float32)])
    r2 = r2[r2 < h2]
    got = (r2[:, None] >= t[None, :]).sum(axis=1)
    assert np.array_equal(got, circle(r2))


def test_set_state_rejects_mismatched_shapes_before_device_work():
    """Simulation.set_state checks shapes before it creates the context (no
    GPU here): the library reads particleCount float4 rows from each array."""
    scene = pbf_amd.config("dam4k")
    sim = pbf_amd.Simulation(scene.params, scene.options)
    n = scene.n
    for x, v in ((np.zeros((n - 1, 4), np.float32), None),
                 (np.zeros((n, 3), np.float32), None),
                 (np.zeros((n, 4), np.float32), np.zeros((n - 1, 4), np.float32)),
                 (np.zeros((n, 4), np.float32), np.zeros((n, 3), np.float32))):
        with pytest.raises(ValueError):
            sim.set_state(x, v)
    assert sim._ctx is None


def test_adjoint_buffers_are_validated():
    """Rollout.forward_raw/backward_raw hand device pointers to the library,
    which reads and writes n float4 rows: anything but a contiguous float32
    (n, 4) GPU tensor is refused first."""
    torch = pytest.importorskip("torch")
    from pbf_amd.adjoint import check_float4
    n = 8
    ok_shape = torch.zeros((n, 4), dtype=torch.float32)
    for t, what in ((torch.zeros((n - 1, 4)), "shape"), (torch.zeros((n, 3)), "shape"),
                    (torch.zeros((n, 4), dtype=torch.float64), "float32"), (ok_shape, "GPU"),
                    (torch.zeros((4, n)).t(), "contiguous")):
        with pytest.raises(ValueError, match=what):
            check_float4(t, n, "x0")
