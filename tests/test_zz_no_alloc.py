"""pbf_run_noalloc (include/pbf/pbf.h): the substeps of one scene with every
intermediate buffer in the library's static scratch, no context and no
allocation, against the context path (pbf_create, pbf_set_state, pbf_run,
pbf_get_state) on the same inputs, bit for bit.  Two calls in a row, the
second with fewer particles, so the second call starts from the first call's
leftovers in the shared scratch."""
import numpy as np
import pytest

import pbf_amd
from pbf_amd import _lib

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _context_run(scene, steps):
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    sim.run(steps)
    st = sim.get_state()
    sim.close()
    return st


def _noalloc_run(scene, steps, v=True):
    n = scene.n
    x = torch.from_numpy(np.ascontiguousarray(scene.x, np.float32)).cuda()
    vv = torch.from_numpy(np.ascontiguousarray(scene.v, np.float32)).cuda() if v else None
    xo = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    vo = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = pbf_amd.lib().pbf_run_noalloc(scene.params, scene.options, x.data_ptr(), None if vv is None else vv.data_ptr(),
                                       n, steps, xo.data_ptr(), vo.data_ptr(), None)
    assert rc == 0, pbf_amd.lib().pbf_last_error(None).decode()
    torch.cuda.synchronize()
    return xo.cpu().numpy(), vo.cpu().numpy()


def test_run_noalloc_equals_the_context_path():
    for name, seed, steps in (("dam16k", 21, 6), ("dam4k", 22, 4)):
        scene = pbf_amd.config(name, seed=seed)
        v = np.zeros_like(scene.v)
        v[:, :3] = np.random.default_rng(seed).normal(0.0, 0.3, (scene.n, 3)).astype(np.float32)
        scene = pbf_amd.Scene(scene.params, scene.options, scene.x, v)
        want = _context_run(scene, steps)
        x, vel = _noalloc_run(scene, steps)
        assert np.array_equal(x.view(np.uint32), want["x"].view(np.uint32)), name
        assert np.array_equal(vel.view(np.uint32), want["v"].view(np.uint32)), name
    # v4 = NULL is zero velocities
    scene = pbf_amd.config("dam4k", seed=23)
    want = _context_run(scene, 2)
    x, vel = _noalloc_run(scene, 2, v=False)
    assert np.array_equal(x.view(np.uint32), want["x"].view(np.uint32))


def test_run_noalloc_rejects_sizes_beyond_its_scratch():
    lib = pbf_amd.lib()
    dummy = torch.zeros((1, 4), dtype=torch.float32, device="cuda")
    big = pbf_amd.dam_break(32, 32, 17, 0.1, (3.2, 1.6, 1.7), seed=1)          # 17,408 > 16,384 particles
    rc = lib.pbf_run_noalloc(big.params, big.options, dummy.data_ptr(), None, big.n, 1, dummy.data_ptr(),
                             dummy.data_ptr(), None)
    assert rc == _lib.PBF_E_INVALID and b"16384" in lib.pbf_last_error(None)
    wide = pbf_amd.dam_break(8, 8, 8, 0.1, (0.8, 0.8, 0.8), seed=1, grid_buf_size=1 << 17)
    rc = lib.pbf_run_noalloc(wide.params, wide.options, dummy.data_ptr(), None, wide.n, 1, dummy.data_ptr(),
                             dummy.data_ptr(), None)
    assert rc == _lib.PBF_E_INVALID and b"65536" in lib.pbf_last_error(None)
