"""Batched scenes (pbf_create_batch): every scene of the batch must equal
the oracle run of that scene alone, bit for bit."""
import numpy as np
import pytest

import pbf_amd
import pbf_oracle

pytestmark = pytest.mark.gpu


def scene_slices(structure, offsets, cell_bases, s, gbs):
    a, b = int(offsets[s]), int(offsets[s + 1])
    n_total = int(offsets[-1])
    n = b - a
    keys = structure["keys"][a:b].astype(np.int64) - cell_bases[s]
    perm = structure["perm"][a:b].astype(np.int64) - a
    cells = structure["cells"][2 * cell_bases[s]: 2 * (cell_bases[s] + gbs)].astype(np.int64)
    cells = np.where(cells == 0xFFFFFFFF, 0xFFFFFFFF, cells - a)
    fr = structure["friends"]
    counts = fr[: 5 * n_total].reshape(5, n_total)[:, a:b]
    idx = fr[5 * n_total:].reshape(5, 50, n_total)[:, :, a:b].astype(np.int64)
    mask = np.arange(50)[None, :, None] < counts[:, None, :]
    idx = np.where(mask, idx - a, 0)
    return keys, perm, cells, counts, idx


def check_batch(scenes, steps, graph=False, quads=None):
    """quads: per scene None or (quads, mask, edge_base)."""
    quads = quads or [None] * len(scenes)
    sim = pbf_amd.BatchSimulation([s.params for s in scenes], [s.options for s in scenes])
    for s, q in enumerate(quads):
        if q is not None:
            sim.set_collision_quads(q[0], q[1], q[2], scene=s)
    sim.set_state(np.concatenate([s.x for s in scenes]), np.concatenate([s.v for s in scenes]))
    if graph:
        sim.set_graph(True)
    sim.run(steps)
    got = sim.get_state()
    st = sim.get_structure()
    off = sim.offsets
    cell_bases = np.concatenate([[0], np.cumsum([int(s.params.gridBufSize) for s in scenes])])
    for s, scene in enumerate(scenes):
        cfg = pbf_oracle.cfg_from_params(scene.params, scene.options)
        if quads[s] is not None:
            pbf_oracle.set_quads(cfg, *quads[s])
        ref = pbf_oracle.OracleSim(cfg)
        ref.set_state(scene.x, scene.v)
        for _ in range(steps):
            ref.step()
        want = ref.get_state()
        rs = ref.get_structure()
        n = scene.n
        keys, perm, cells, counts, idx = scene_slices(st, off, cell_bases, s, int(scene.params.gridBufSize))
        assert np.array_equal(keys, rs["keys"]), s
        assert np.array_equal(perm, rs["perm"]), s
        assert np.array_equal(cells, rs["cells"].astype(np.int64)), s
        rf = rs["friends"]
        assert np.array_equal(counts, rf[: 5 * n].reshape(5, n)), s
        assert np.array_equal(idx, rf[5 * n:].reshape(5, 50, n).astype(np.int64)), s
        a, b = int(off[s]), int(off[s + 1])
        for k in ("x", "v", "xs", "omega"):
            assert np.array_equal(got[k][a:b].view(np.uint32), want[k].view(np.uint32)), (s, k)
        ref.close()
    sim.close()


def test_batch_matches_single_scenes():
    scenes = pbf_amd.ensemble(6, seed=1, n_min=500, n_max=6000)
    check_batch(scenes, 3)


def test_batch_graph_replay():
    scenes = pbf_amd.ensemble(5, seed=2, n_min=800, n_max=5000)
    check_batch(scenes, 4, graph=True)


def test_batch_per_scene_gravity():
    scenes = pbf_amd.ensemble(3, seed=3, n_min=1000, n_max=3000)
    for s, g in zip(scenes, ((0.0, -9.8, 0.0), (3.0, -9.0, 1.5), (-2.0, -7.0, -6.2))):
        s.options = pbf_amd.default_options(gravity=g)
    check_batch(scenes, 2)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_per_scene_quads(graph):
    from _quads import box_quads
    scenes = pbf_amd.ensemble(3, seed=4, n_min=1000, n_max=3000)
    q0, m0 = box_quads(scenes[0].params, mask_seed=1)
    q2, m2 = box_quads(scenes[2].params, plate_x=0.3, mask_seed=2)
    check_batch(scenes, 6, graph=graph, quads=[(q0, m0, 0.002), None, (q2[1:], m2, 0.0)])


def test_full_ensemble_equals_its_scenes():
    # the bench's 64-scene ensemble (about 2.3M particles): the batch equals
    # each scene stepped alone, bit for bit
    scenes = pbf_amd.ensemble(64, seed=0)
    steps = 3
    batch = pbf_amd.BatchSimulation([s.params for s in scenes], [s.options for s in scenes])
    batch.set_state(np.concatenate([s.x for s in scenes]), np.concatenate([s.v for s in scenes]))
    batch.set_graph(True)
    batch.run(steps)
    got = batch.get_state()
    off = batch.offsets
    batch.close()
    for k, s in enumerate(scenes):
        sim = pbf_amd.Simulation(s.params, s.options)
        sim.set_state(s.x, s.v)
        sim.run(steps)
        one = sim.get_state()
        sim.close()
        a, b = int(off[k]), int(off[k + 1])
        assert np.array_equal(got["x"][a:b].view(np.uint32), one["x"].view(np.uint32)), k
        assert np.array_equal(got["v"][a:b].view(np.uint32), one["v"].view(np.uint32)), k
