"""CPU checks of the adjoint's checker (oracle/pbf_adjoint_ref.cpp).

The frozen-structure rollout is pinned three ways before it is used to judge
the GPU adjoint:
  * its binary32 instantiation reproduces the binary32 oracle (the
    restatement of the reference kernels) bit for bit, so it differentiates
    the same function;
  * its FP64 dual numbers agree with FP64 central differences of the whole
    rollout (the tangent is the derivative);
  * its binary32-primal dual numbers agree with FP64 central differences of
    every substep chained along the binary32 trajectory.
"""
import numpy as np
import pytest

import pbf_amd


@pytest.fixture(scope="module")
def frozen(oracle):
    scene = pbf_amd.scenes.inverse_problem(8, 8, 8, seed=3)
    cfg = oracle.cfg_from_params(scene.params, scene.options)
    return oracle.FrozenRollout(cfg, scene.x, scene.v, 8)


# a ramp over the floor (normal up and towards -x) and a plate facing -x at
# the far side of the block, both with a half-solid random mask
QUADS = [((0.0, 0.03, 0.0), (0.0, 0.0, 0.8), (0.8, 0.1, 0.0), 0, 16),
         ((0.38, 0.0, 0.0), (0.0, 0.0, 0.8), (0.0, 0.8, 0.0), 16, 16)]


def quad_mask(seed=0):
    return (np.random.default_rng(seed).random((32, 16)) < 0.5).astype(np.uint32)


@pytest.fixture(scope="module")
def frozen_quads(oracle):
    scene = pbf_amd.scenes.inverse_problem(8, 8, 8, seed=3)
    cfg = oracle.cfg_from_params(scene.params, scene.options)
    oracle.set_quads(cfg, QUADS, quad_mask(), 0.004)
    return oracle.FrozenRollout(cfg, scene.x, scene.v, 8)


def _direction(fr, what, seed):
    rng = np.random.default_rng(seed)
    dx = dv = dc = None
    if what == "v0":
        dv = rng.normal(0, 1, (fr.n, 3))
    elif what == "x0":
        dx = rng.normal(0, 1e-3, (fr.n, 3))
    else:
        dc = np.zeros(5)
        if what == "gravity":
            dc[2:] = rng.normal(0, 1, 3)
        else:
            dc[{"viscosity": 0, "vorticity": 1}[what]] = 1.0
    return dx, dv, dc


def test_binary32_instantiation_is_the_oracle(frozen):
    x32, v32, _, _ = frozen.jvp32()
    assert np.array_equal(x32, frozen.xT32)
    assert np.array_equal(v32, frozen.vT32)


def test_fp64_rollout_tracks_binary32_over_a_short_horizon(frozen):
    x64, v64 = frozen.rollout64()
    assert np.abs(x64 - frozen.xT32).max() < 1e-4   # 1e-3 h at h = 0.1
    assert np.isfinite(v64).all()


@pytest.mark.parametrize("what", ["v0", "x0", "viscosity", "vorticity", "gravity"])
def test_fp64_jvp_equals_central_differences(frozen, what):
    dx, dv, dc = _direction(frozen, what, 1)
    tgt = frozen.xT32 + np.random.default_rng(0).normal(0, 0.01, frozen.xT32.shape)

    def loss(xT):
        return np.mean(np.sum((xT - tgt) ** 2, axis=1))

    xb, _, dxT, _ = frozen.jvp64(dx0=dx, dv0=dv, dcoef=dc)
    jvp = np.sum(2.0 * (xb - tgt) * dxT) / frozen.n
    x0 = frozen.x0.astype(np.float64)
    v0 = frozen.v0.astype(np.float64)
    c0 = frozen.coef0.astype(np.float64)
    # steps: 1e-6 on the state and gravity; the XSPH and vorticity factors are
    # ~1e-6 and ~1e-7 themselves, so a 1e-4 relative step keeps the round-off
    # of the difference below its truncation error
    e = {"viscosity": 1e-4 * abs(c0[0]), "vorticity": 1e-4 * abs(c0[1])}.get(what, 1e-6)

    def L(k):
        return loss(frozen.rollout64(x0 + (0 if dx is None else k * e * dx), v0 + (0 if dv is None else k * e * dv),
                                     c0 + (0 if dc is None else k * e * dc))[0])

    # fourth-order central difference
    fd = (L(-2.0) - 8.0 * L(-1.0) + 8.0 * L(1.0) - L(2.0)) / (12.0 * e)
    assert abs(fd - jvp) <= 1e-5 * abs(jvp), (fd, jvp)


@pytest.mark.parametrize("what", ["v0", "viscosity", "vorticity", "gravity"])
def test_binary32_jvp_equals_chained_fp64_differences(frozen, what):
    _, dv, dc = _direction(frozen, what, 2)
    _, _, dx32, dv32 = frozen.jvp32(dv0=dv, dcoef=dc)
    dxf, dvf = frozen.tangent_fd(dv0=dv, dcoef=dc, delta=1e-6)
    assert np.abs(dx32 - dxf).max() <= 1e-4 * np.abs(dx32).max()
    assert np.abs(dv32 - dvf).max() <= 1e-4 * np.abs(dv32).max()


def test_frozen_structure_records_the_walls(frozen):
    # the block starts in the box corner with random velocities: the three
    # low planes fire within a few substeps, the high ones (far away) do not
    lo = frozen.hits & 0b010101
    hi = frozen.hits & 0b101010
    assert lo.any()
    assert not hi.any()


def test_frozen_structure_records_the_quads(frozen_quads):
    q = frozen_quads.hits >> 6
    assert (q & 1).any() and (q & 2).any()
    assert not (q >> 2).any()


def test_binary32_instantiation_with_quads_is_the_oracle(frozen_quads):
    x32, v32, _, _ = frozen_quads.jvp32()
    assert np.array_equal(x32, frozen_quads.xT32)
    assert np.array_equal(v32, frozen_quads.vT32)


@pytest.mark.parametrize("what", ["v0", "gravity"])
def test_fp64_jvp_with_quads_equals_central_differences(frozen_quads, what):
    fr = frozen_quads
    dx, dv, dc = _direction(fr, what, 3)
    tgt = fr.xT32 + np.random.default_rng(1).normal(0, 0.01, fr.xT32.shape)

    def loss(xT):
        return np.mean(np.sum((xT - tgt) ** 2, axis=1))

    xb, _, dxT, _ = fr.jvp64(dx0=dx, dv0=dv, dcoef=dc)
    jvp = np.sum(2.0 * (xb - tgt) * dxT) / fr.n
    x0, v0, c0 = fr.x0.astype(np.float64), fr.v0.astype(np.float64), fr.coef0.astype(np.float64)
    e = 1e-6

    def L(k):
        return loss(fr.rollout64(x0, v0 + (0 if dv is None else k * e * dv), c0 + (0 if dc is None else k * e * dc))[0])

    fd = (L(-2.0) - 8.0 * L(-1.0) + 8.0 * L(1.0) - L(2.0)) / (12.0 * e)
    assert abs(fd - jvp) <= 1e-5 * abs(jvp), (fd, jvp)


def test_binary32_jvp_with_quads_equals_chained_fp64_differences(frozen_quads):
    _, dv, _ = _direction(frozen_quads, "v0", 4)
    _, _, dx32, dv32 = frozen_quads.jvp32(dv0=dv)
    dxf, dvf = frozen_quads.tangent_fd(dv0=dv, delta=1e-6)
    assert np.abs(dx32 - dxf).max() <= 1e-4 * np.abs(dx32).max()
    assert np.abs(dv32 - dvf).max() <= 1e-4 * np.abs(dv32).max()
