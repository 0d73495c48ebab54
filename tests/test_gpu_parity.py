"""GPU parity: the HIP substep (through the C ABI) against the CPU oracle.

Integer structure (cell keys, sort permutation, cells, friends lists) must
match bit for bit.  Floating-point state must agree within 1e-5*h in
positions (1e-5*h/dt in velocities); both sides evaluate the same binary32
operations in the same order, fusing multiply-adds at the same points only,
so the state is also checked for bitwise equality.
"""
import numpy as np
import pytest

import pbf_amd
import pbf_oracle

pytestmark = pytest.mark.gpu


def make_pair(scene):
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    ref = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(scene.params, scene.options))
    ref.set_state(scene.x, scene.v)
    return sim, ref


def advance(sim, ref, steps=1, pauses=None):
    for s in range(steps):
        pause = bool(pauses[s]) if pauses is not None else False
        sim.bPauseSim = pause
        sim.Step()
        ref.step(pause)


def check_structure(sim, ref, friends=True):
    a = sim.get_structure(friends=friends)
    b = ref.get_structure(friends=friends)
    assert np.array_equal(a["keys"], b["keys"]), "cell keys differ"
    assert np.array_equal(a["perm"], b["perm"]), "sort permutation differs"
    assert np.array_equal(a["cells"], b["cells"]), "cell start/end differ"
    if friends:
        assert np.array_equal(a["friends"], b["friends"]), "friends lists differ"
    return a


def check_state(sim, ref, scene, bitwise=True):
    """All four components of x (w = |v|, the render VBO), x* (w = density,
    the render texture), v and omega (w = |omega|), then rho and lambda."""
    a = sim.get_state()
    b = ref.get_state()
    h = float(scene.params.h)
    dt = float(scene.params.timeStep)
    for k, tol in (("x", 1e-5 * h), ("xs", 1e-5 * h), ("v", 1e-5 * h / dt), ("omega", None)):
        da, db = a[k], b[k]
        assert np.array_equal(np.isnan(da), np.isnan(db)), f"{k}: NaN pattern differs"
        fin = ~np.isnan(db)
        err = float(np.max(np.abs(da[fin] - db[fin]))) if fin.any() else 0.0
        if tol is not None and k != "xs":   # x*.w holds the density, not a position
            assert err <= tol, f"{k}: max |diff| {err:.3e} > {tol:.3e}"
        elif k == "xs":
            e3 = float(np.max(np.abs(da[:, :3][fin[:, :3]] - db[:, :3][fin[:, :3]]))) if fin.any() else 0.0
            assert e3 <= tol, f"{k}: max |diff| {e3:.3e} > {tol:.3e}"
        if bitwise:
            assert np.array_equal(da.view(np.uint32)[fin], db.view(np.uint32)[fin]), f"{k}: not bitwise equal (max {err:.3e})"
    for k in ("rho", "lambda"):
        if bitwise:
            fin = ~np.isnan(b[k])
            assert np.array_equal(a[k].view(np.uint32)[fin], b[k].view(np.uint32)[fin]), f"{k} not bitwise equal"
    return a, b


@pytest.mark.parametrize("name", ["dam4k", "dam16k"])
def test_one_step(name):
    scene = pbf_amd.config(name, seed=3)
    sim, ref = make_pair(scene)
    advance(sim, ref, 1)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    flags, dropped = sim.error_flags()
    assert flags == 0 and dropped == 0
    sim.close()


def test_ten_steps_dam16k():
    scene = pbf_amd.config("dam16k", seed=5)
    sim, ref = make_pair(scene)
    advance(sim, ref, 10)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


def test_ragged_count():
    # 3,000 particles: not a multiple of the 256-thread blocks or 2,048-key sort tiles
    scene = pbf_amd.dam_break(15, 20, 10, 0.1, (1.6, 2.4, 1.6), seed=7)
    sim, ref = make_pair(scene)
    advance(sim, ref, 3)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


@pytest.mark.parametrize("gbs", [1, 64, 97, 4093])
def test_hash_folding(gbs):
    # a small GRID_BUF_SIZE folds distinct cells onto one hash: neighbour
    # cells are visited once per alias, so friends lists carry duplicates
    scene = pbf_amd.dam_break(12, 12, 12, 0.1, (1.2, 1.2, 1.2), seed=11, grid_buf_size=gbs)
    if gbs == 1:
        # every particle shares one cell: 27 visits of all particles, keep it tiny
        scene = pbf_amd.dam_break(4, 4, 4, 0.1, (0.4, 0.4, 0.4), seed=11, grid_buf_size=gbs)
    sim, ref = make_pair(scene)
    advance(sim, ref, 2)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


@pytest.mark.parametrize("n", [1, 2, 65])
def test_tiny_counts(n):
    x = pbf_amd.dam_break_block(n, 1, 1, 0.1, seed=2)
    x[:, 1] += 0.3
    p = pbf_amd.physical_params(n, 0.1, (0, 0, 0), (8.0, 1.0, 1.0))
    scene = pbf_amd.Scene(p, pbf_amd.default_options(), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    advance(sim, ref, 3)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_scenes(seed):
    """Scenes drawn at random (count, h, box, grid size with and without
    folding, solver iterations, gravity vector, walls, initial velocities,
    particles outside the box, pause pattern) give the oracle's structure and
    state bit for bit over a few substeps."""
    rng = np.random.default_rng(7000 + seed)
    n = int(rng.integers(40, 3000))
    h = float(rng.choice([0.05, 0.1, 0.25]))
    extent = rng.uniform(8 * h, 30 * h, 3)
    lo = rng.uniform(-2.0, 2.0, 3)
    gbs = [None, 64, 1000, 4096, 12345, 1 << 16][seed]
    iters = int(rng.integers(1, 7))
    p = pbf_amd.physical_params(n, h, lo, lo + extent, iters=iters, grid_buf_size=gbs,
                                xsph_c=float(rng.uniform(0.0, 0.05)), vort_eps=float(rng.uniform(0.0, 1e-3)))
    g = rng.normal(0.0, 5.0, 3) if seed % 2 else None
    opts = pbf_amd.default_options(walls=bool(seed % 3), gravity=g)
    # a block at about twice the rest density in one corner and a sparse
    # cloud, a few particles outside the box
    dense = n * 2 // 3
    side = np.minimum(extent, (dense / 16.0) ** (1.0 / 3.0) * h)
    x = np.zeros((n, 4), np.float32)
    x[:dense, :3] = lo + rng.uniform(0.0, 1.0, (dense, 3)) * side
    x[dense:, :3] = lo + rng.uniform(-0.1, 1.1, (n - dense, 3)) * extent
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0.0, 0.5, (n, 3))
    scene = pbf_amd.Scene(p, opts, x, v)
    sim, ref = make_pair(scene)
    advance(sim, ref, 4, pauses=list(rng.integers(0, 2, 4)))
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


# pairs whose |r|^2 with fmaf and unfused lie on either side of h^2 = 0.1f^2
_STRADDLING_PAIRS = [
    ((0.2739469110965729, 0.3455261290073395, 0.5627720355987549),
     (0.2296944111585617, 0.379925400018692, 0.4799565374851227)),
    ((0.4090348780155182, 0.3165309727191925, 0.20136530697345734),
     (0.42549607157707214, 0.22700658440589905, 0.24277077615261078)),
    ((0.4496609568595886, 0.23794008791446686, 0.21546831727027893),
     (0.46433621644973755, 0.1931574046611786, 0.12726883590221405)),
]


@pytest.mark.parametrize("k", range(len(_STRADDLING_PAIRS)))
def test_friend_test_at_the_fused_distance(k):
    """A paused substep from rest keeps positions exact, so the friends build
    sees each pair as placed: its friend test agrees with the oracle on pairs
    whose fused and unfused |r|^2 straddle h^2."""
    a, b = _STRADDLING_PAIRS[k]
    x = np.zeros((2, 4), np.float32)
    x[0, :3], x[1, :3] = a, b
    p = pbf_amd.physical_params(2, 0.1, (0, 0, 0), (1.0, 1.0, 1.0))
    scene = pbf_amd.Scene(p, pbf_amd.default_options(walls=False), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    advance(sim, ref, 1, pauses=[1])
    st = check_structure(sim, ref)
    sim.close()
    fr = st["friends"]
    assert int(fr[:10].sum()) == [0, 2, 0][k]   # pairs 0 and 2 are not friends, pair 1 is


def test_lattice_on_cell_boundaries():
    """A lattice at spacing h/2 through the origin (h = 2^-3, every coordinate
    exact, half of them on cell faces, negative ones in the cell of
    truncation toward zero): a paused substep from rest keeps them there, so
    the cell keys, the neighbour-cell cull and the friend tests all meet
    exact boundary cases; structure and state equal the oracle's."""
    h = 0.125
    g = np.arange(-6, 7) * (h / 2)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.zeros((X.size, 4), np.float32)
    x[:, 0], x[:, 1], x[:, 2] = X.ravel(), Y.ravel(), Z.ravel()
    p = pbf_amd.physical_params(len(x), h, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    scene = pbf_amd.Scene(p, pbf_amd.default_options(walls=False), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    for _ in range(2):   # the first substep's build sees the lattice as placed
        advance(sim, ref, 1, pauses=[1])
        check_structure(sim, ref)
        check_state(sim, ref, scene)
    sim.close()


def test_pause_semantics():
    scene = pbf_amd.config("dam4k", seed=9)
    sim, ref = make_pair(scene)
    advance(sim, ref, 5, pauses=[1, 0, 1, 1, 0])
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


def test_friend_overflow_flag():
    # 400 particles within 0.2h of each other: every pair lands in circle 0,
    # which keeps the first 50 in scan order and drops the rest
    rng = np.random.default_rng(0)
    n = 400
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = (0.55 + 0.02 * rng.random((n, 3))).astype(np.float32)
    p = pbf_amd.physical_params(n, 0.1, (0, 0, 0), (1.0, 1.0, 1.0), iters=1)
    scene = pbf_amd.Scene(p, pbf_amd.default_options(), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    advance(sim, ref, 1)
    st = check_structure(sim, ref)
    flags, dropped = sim.error_flags(reset=True)
    counts = st["friends"][: 5 * n].reshape(5, n)
    assert flags & pbf_amd.PBF_FLAG_FRIENDS_OVERFLOW
    assert int(counts[0].max()) == 50
    # every particle sees n-1 friends in range; all beyond 50 per circle drop
    assert dropped == int(n * (n - 1) - counts.sum())
    assert sim.error_flags() == (0, 0)
    # the event log names every particle, with its drop count and circle counts
    log, lost = sim.read_log()
    assert lost == 0 and len(log) == n
    assert sorted(m["a"] for m in log) == list(range(n))
    assert all(m["event"] == "friends_overflow" and m["b"] == n - 1 - 50 and m["c"] == 50 for m in log)
    assert sim.read_log() == ([], 0)
    sim.close()


def test_non_finite_particles_match_oracle():
    # a NaN and an infinite particle land in friend lists (as in the reference,
    # whose r^2 >= h^2 test lets NaN through); the sweeps skip their pairs like
    # the oracle's branches, so the finite particles stay finite and bitwise
    # equal to the oracle
    scene = pbf_amd.config("dam4k", seed=3)
    x = scene.x.copy()
    x[7, 0] = np.nan
    x[100, 2] = np.inf
    scene = pbf_amd.Scene(scene.params, scene.options, x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    advance(sim, ref, 3)
    check_structure(sim, ref)
    a, b = check_state(sim, ref, scene)
    assert np.isfinite(a["x"][:, :3]).all(axis=1).sum() >= scene.n - 2
    sim.close()


def test_event_log_nan_and_wrap():
    # a non-finite position is logged once for its workgroup; the ring keeps the
    # last 1024 messages and counts the overwritten ones
    scene = pbf_amd.config("dam4k", seed=2)
    x = scene.x.copy()
    x[[5, 9], 1] = np.nan
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(x, scene.v)
    sim.Step()
    log, lost = sim.read_log()
    nan = [m for m in log if m["event"] == "nan"]
    assert lost == 0 and len(nan) == 1 and nan[0]["a"] == 5 and nan[0]["b"] == 2 and nan[0]["c"] == 0
    assert sim.error_flags()[0] & pbf_amd.PBF_FLAG_NAN
    sim.close()
    # 1,200 overflowing particles in one step: 176 messages are lost
    rng = np.random.default_rng(1)
    n = 1200
    xo = np.zeros((n, 4), np.float32)
    xo[:, :3] = (0.55 + 0.02 * rng.random((n, 3))).astype(np.float32)
    p = pbf_amd.physical_params(n, 0.1, (0, 0, 0), (1.0, 1.0, 1.0), iters=1)
    sim = pbf_amd.Simulation(p, pbf_amd.default_options())
    sim.set_state(xo, np.zeros_like(xo))
    sim.Step()
    first, lost = sim.read_log(cap=1000)
    rest, lost2 = sim.read_log()
    assert lost == n - 1024 and lost2 == 0 and len(first) == 1000 and len(rest) == 24
    assert len({m["a"] for m in first + rest}) == 1024
    sim.close()


def test_coincident_particles():
    # pairs of particles at exactly one point: finite (no 0/0) and bitwise
    # equal to the oracle
    scene = pbf_amd.config("dam4k", seed=11)
    x = scene.x.copy()
    x[1::2, :3] = x[0::2, :3]
    scene = pbf_amd.Scene(scene.params, scene.options, x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    advance(sim, ref, 3)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    assert np.isfinite(sim.get_state()["x"]).all()
    sim.close()


def test_collision_quads():
    # a ramp over the floor and a plate in the path of the dam break, with a
    # random mask: bitwise equal to the oracle, and the quads do move points
    from _quads import box_quads
    scene = pbf_amd.config("dam4k", seed=12)
    quads, mask = box_quads(scene.params)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_collision_quads(quads, mask, 0.003)
    sim.set_state(scene.x, scene.v)
    cfg = pbf_oracle.set_quads(pbf_oracle.cfg_from_params(scene.params, scene.options), quads, mask, 0.003)
    ref = pbf_oracle.OracleSim(cfg)
    ref.set_state(scene.x, scene.v)
    advance(sim, ref, 24)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    plain = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(scene.params, scene.options))
    plain.set_state(scene.x, scene.v)
    for _ in range(24):
        plain.step()
    assert not np.array_equal(plain.get_state()["x"], ref.get_state()["x"])
    # removing the quads restores the plain substep
    sim.set_collision_quads([])
    sim.set_state(scene.x, scene.v)
    sim.run(24)
    assert np.array_equal(sim.get_state()["x"], plain.get_state()["x"])
    sim.close()


def test_collision_quads_graph_replay():
    from _quads import box_quads
    scene = pbf_amd.config("dam4k", seed=13)
    quads, mask = box_quads(scene.params, plate_x=0.4)
    out = []
    for graph in (False, True):
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_graph(graph)
        sim.set_state(scene.x, scene.v)
        sim.run(2)
        sim.set_collision_quads(quads, mask)   # after capture: the graph reads the new table
        sim.run(10)
        out.append(sim.get_state()["x"])
        sim.close()
    assert np.array_equal(out[0], out[1])


def test_collision_quads_are_validated():
    scene = pbf_amd.config("dam4k", seed=1)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    ok = ((0, 0, 0), (1, 0, 0), (0, 0, 1), 0, 1)
    for bad in ([((0, 0, 0), (1, 0, 0), (2, 0, 0), 0, 1)],          # parallel edges
                [((0, 0, 0), (float("nan"), 0, 0), (0, 0, 1), 0, 1)],
                [((0, 0, 0), (1, 0, 0), (0, 0, 1), 1 << 25, 1)]):
        with pytest.raises(pbf_amd.PbfError) as e:
            sim.set_collision_quads(bad)
        assert e.value.code == pbf_amd._lib.PBF_E_INVALID
    with pytest.raises(ValueError):
        sim.set_collision_quads([ok] * 27)
    with pytest.raises(pbf_amd.PbfError):
        sim.set_collision_quads([ok], np.zeros((0, 4), np.uint32))
    sim.set_collision_quads([ok] * 26)
    sim.close()


def test_update_params_matches_a_fresh_oracle():
    # a reloaded scenario: after 4 steps the context takes new parameters; the
    # next 4 steps equal an oracle started from the same state with them
    scene = pbf_amd.config("dam4k", seed=6)
    sim, ref = make_pair(scene)
    advance(sim, ref, 4)
    p2 = scene.params.copy()
    p2.garvity = 4.0
    p2.viscosityFactor *= 3.0
    p2.h = 0.105
    p2.timeStep = 0.007
    pbf_amd.lib().pbf_params_finalize(p2)
    sim.update_params(p2)
    st = ref.get_state()
    ref2 = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(p2, scene.options))
    ref2.set_state(st["x"], st["v"])
    advance(sim, ref2, 4)
    check_structure(sim, ref2)
    check_state(sim, ref2, pbf_amd.Scene(p2, scene.options, scene.x, scene.v))
    p3 = p2.copy()
    p3.particleCount += 1
    with pytest.raises(pbf_amd.PbfError) as e:
        sim.update_params(p3)
    assert e.value.code == pbf_amd._lib.PBF_E_INVALID
    sim.close()


def test_deterministic_repeat():
    scene = pbf_amd.config("dam16k", seed=4)
    outs = []
    for _ in range(2):
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_state(scene.x, scene.v)
        sim.run(8)
        outs.append(sim.get_state())
        sim.close()
    for k in ("x", "v", "xs", "omega"):
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), k


def _traces(state, rest_density):
    rho = state["rho"].astype(np.float64)
    v = state["v"][:, :3].astype(np.float64)
    return float(np.mean(np.abs(rho / rest_density - 1.0))), float(0.5 * np.sum(v * v))


def test_trace_200_steps_dam16k():
    """Density-error and kinetic-energy traces over 200 substeps within 1%."""
    scene = pbf_amd.config("dam16k", seed=0)
    sim, ref = make_pair(scene)
    rd = float(scene.params.restDensity)
    worst = 0.0
    for s in range(200):
        advance(sim, ref, 1)
        if s % 10 == 9:
            da, ka = _traces(sim.get_state(), rd)
            db, kb = _traces(ref.get_state(), rd)
            worst = max(worst, abs(da - db) / max(abs(db), 1e-30), abs(ka - kb) / max(abs(kb), 1e-30))
    assert worst <= 0.01, worst
    flags, _ = sim.error_flags()
    assert not (flags & pbf_amd.PBF_FLAG_SORT_TIMEOUT)
    sim.close()


def test_one_step_dam1m():
    """Full-size workload: structure bit-exact, state within tolerance."""
    scene = pbf_amd.config("dam1m", seed=0)
    sim, ref = make_pair(scene)
    advance(sim, ref, 1)
    a = check_structure(sim, ref)
    n = scene.n
    keys = a["keys"]
    assert np.all(keys[1:] >= keys[:-1])
    assert np.array_equal(np.sort(a["perm"]), np.arange(n, dtype=np.uint32))
    check_state(sim, ref, scene)
    flags, dropped = sim.error_flags()
    assert flags == 0 and dropped == 0
    sim.close()


def test_dam1m_long_run_properties():
    """The full 1,048,576-particle workload over its 500 substeps (past the
    point where the oracle is practical), through the regime where friends
    drop at the 50-per-circle cap: every particle stays finite and inside the
    box, graph replay equals eager launches bit for bit, and a second run
    repeats the first."""
    scene = pbf_amd.config("dam1m", seed=0)
    p = scene.params
    runs = []
    for graph in (False, True, True):
        sim = pbf_amd.Simulation(p, scene.options)
        sim.set_graph(graph)
        sim.set_state(scene.x, scene.v)
        sim.run(500)
        flags, dropped = sim.error_flags()
        assert not (flags & (pbf_amd.PBF_FLAG_SORT_TIMEOUT | pbf_amd.PBF_FLAG_NAN)), flags
        assert not (flags & 8), "internal bounds check"
        st = sim.get_state()
        runs.append(st)
        sim.close()
    x = runs[0]["x"][:, :3]
    assert np.isfinite(x).all()
    lo = np.array([p.xMin, p.yMin, p.zMin], np.float32)
    hi = np.array([p.xMax, p.yMax, p.zMax], np.float32)
    # A particle the walls place on a plane where the noise edge
    # (3 + sum sin) 0.03 h rounds to 0 sits exactly on the box face; if it
    # keeps moving out, the force-zone test |dp| > edge + nv of the next
    # substep (compute_delta.cl:29-31) is an equality in exact arithmetic and
    # can tip by one rounding, and a particle outside is never caught again.
    # The oracle does the same: in this workload one particle leaves through
    # x = 12.8 at substep 168 and slides on along the floor.
    over = np.maximum(lo - x, x - hi).max(axis=1)
    out = int((over > 0).sum())
    print(f"outside the box after 500 substeps: {out} particles, by at most {float(over.max()):.3e}")
    assert out <= 16, out
    assert np.array_equal(runs[0]["x"].view(np.uint32), runs[1]["x"].view(np.uint32))
    assert np.array_equal(runs[1]["x"].view(np.uint32), runs[2]["x"].view(np.uint32))
    assert np.array_equal(runs[0]["v"].view(np.uint32), runs[2]["v"].view(np.uint32))
    # the fluid has moved: kinetic energy is nonzero and the block has spread in x
    assert float(np.abs(runs[0]["v"][:, :3]).max()) > 0.1
    assert float(x[:, 0].max()) > float(scene.x[:, 0].max()) + 0.5


def test_dam1m_structure_in_the_drop_regime():
    """The 1,048,576-particle dam break from substep 100 on reaches substeps
    where circles hold 50 friends and later ones drop
    (build_friends_list.cl:79-83).  The first such substep, started by the
    oracle from the HIP state just before it, gives the same cell keys,
    permutation, cells and friends lists (with the same drops) bit for bit,
    and the same state."""
    scene = pbf_amd.config("dam1m", seed=0)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    sim.set_graph(True)
    sim.run(100)
    sim.error_flags(reset=True)
    sim.set_graph(False)
    step = None
    for s in range(100, 400):
        before = sim.get_state()
        sim.Step()
        flags, dropped = sim.error_flags(reset=True)
        if dropped:
            step = s
            break
    assert step is not None, "no friends dropped in substeps 100..399"
    print(f"first substep with drops: {step} ({dropped} friends dropped)")
    ref = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(scene.params, scene.options))
    ref.set_state(before["x"], before["v"])
    ref.step()
    st = check_structure(sim, ref)
    counts = st["friends"][: 5 * scene.n].reshape(5, scene.n)
    assert int(counts.max()) == 50 and flags & pbf_amd.PBF_FLAG_FRIENDS_OVERFLOW
    check_state(sim, ref, scene)
    ref.close()
    sim.close()


def test_profiling_trackers():
    scene = pbf_amd.config("dam4k", seed=1)
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    sim.set_profiling(True)
    sim.run(3)
    t = sim.timings()
    # the first computeScaling runs inside buildFriendsList
    for name in ("predictPositions", "onesweep_0", "sortParticles", "buildFriendsList",
                 "computeScaling_1", "computeDelta_3", "applyViscosity", "applyVorticity", "eventTare"):
        assert name in t and t[name]["calls"] == 3 and t[name]["last_ms"] > 0.0, name
    # the empty kernel's duration is the fixed part of every launch's
    assert t["eventTare"]["sum_ms"] < t["buildFriendsList"]["sum_ms"]
    sim.close()


def test_set_state_twice_resets_cells():
    """A second set_state on the same context starts from clean cells."""
    scene = pbf_amd.config("dam4k", seed=12)
    sim, ref = make_pair(scene)
    advance(sim, ref, 3)
    other = pbf_amd.config("dam4k", seed=13)
    sim.set_state(other.x, other.v)
    ref2 = pbf_oracle.OracleSim(pbf_oracle.cfg_from_params(other.params, other.options))
    ref2.set_state(other.x, other.v)
    advance(sim, ref2, 2)
    check_structure(sim, ref2)
    check_state(sim, ref2, other)
    sim.close()


def test_graph_replay_matches_eager():
    """Substeps replayed from captured HIP graphs equal eager launches,
    checked on every replay with eager host work in between."""
    scene = pbf_amd.config("dam16k", seed=6)
    eager = pbf_amd.Simulation(scene.params, scene.options)
    eager.set_state(scene.x, scene.v)
    graph = pbf_amd.Simulation(scene.params, scene.options)
    graph.set_state(scene.x, scene.v)
    graph.set_graph(True)
    for s in range(6):
        eager.Step()
        graph.Step()
        a, b = eager.get_state(), graph.get_state()
        for k in ("x", "v", "xs", "omega"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (s, k)
    # pause variants are captured as their own graphs
    eager.bPauseSim = graph.bPauseSim = True
    eager.Step(); graph.Step()
    a, b = eager.get_state(), graph.get_state()
    assert np.array_equal(a["x"].view(np.uint32), b["x"].view(np.uint32))
    eager.close()
    graph.close()


def test_grouped_graph_runs_match_eager():
    """pbf_run replays groups of substeps as one graph each (with single-
    substep graphs for the rest): every run equals eager substeps bit for
    bit, across both buffer rotations, after a new state and when paused."""
    scene = pbf_amd.config("dam16k", seed=11)
    eager = pbf_amd.Simulation(scene.params, scene.options)
    graph = pbf_amd.Simulation(scene.params, scene.options)
    graph.set_graph(True)

    def check(tag):
        a, b = eager.get_state(), graph.get_state()
        for k in ("x", "v", "xs", "omega", "rho", "lambda"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (tag, k)

    for round_ in range(2):   # the second round replays the graphs the first captured
        eager.set_state(scene.x, scene.v)
        graph.set_state(scene.x, scene.v)
        eager.bPauseSim = graph.bPauseSim = False
        for k in (3, 4, 9, 1, 8):
            for _ in range(k):
                eager.Step()
            graph.run(k)
            check((round_, k))
        eager.bPauseSim = graph.bPauseSim = True
        for _ in range(5):
            eager.Step()
        graph.run(5)
        check((round_, "paused"))
    assert eager.error_flags() == graph.error_flags()
    eager.close()
    graph.close()


def test_workgroup_sizes_do_not_change_results():
    scene = pbf_amd.config("dam16k", seed=7)
    outs = []
    for size in (64, 128, 256, 512):
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_state(scene.x, scene.v)
        sim.set_workgroup_sizes({f: size for f in pbf_amd._lib.FAMILIES})
        sim.run(4)
        outs.append(sim.get_state())
        sim.close()
    for o in outs[1:]:
        for k in ("x", "v", "xs", "omega"):
            assert np.array_equal(outs[0][k].view(np.uint32), o[k].view(np.uint32)), k


@pytest.mark.parametrize("tiles", [2, 4, 0])
def test_tiles_per_workgroup_do_not_change_results(tiles):
    """Strided (2, 4 tiles per workgroup) and persistent (0) launches of the
    friends build and the sweeps give the one-tile-per-workgroup results bit
    for bit, for one scene and for a batch of uneven scenes."""
    fams = pbf_amd._lib.FAMILIES
    scene = pbf_amd.config("dam16k", seed=9)
    scenes = pbf_amd.ensemble(5, seed=4, n_min=700, n_max=20000)
    for make, x0, v0 in (
        (lambda: pbf_amd.Simulation(scene.params, scene.options), scene.x, scene.v),
        (lambda: pbf_amd.BatchSimulation([s.params for s in scenes], [s.options for s in scenes]),
         np.concatenate([s.x for s in scenes]), np.concatenate([s.v for s in scenes])),
    ):
        outs = []
        for t in (1, tiles):
            sim = make()
            sim.set_state(x0, v0)
            sim.set_workgroup_sizes({f: 64 for f in fams})   # many tiles per launch
            with pytest.raises(pbf_amd.PbfError):
                sim.set_tiles_per_workgroup({f: 65 for f in fams})   # at most 64
            sim.set_tiles_per_workgroup({f: t for f in fams})
            assert sim.stats()["tiles_per_workgroup"] == {f: t for f in fams}
            sim.run(4)
            outs.append(sim.get_state())
            sim.close()
        for k in ("x", "v", "xs", "omega"):
            assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), (tiles, k)


def test_strided_tiles_under_graph_replay():
    """Strided launches (4 tiles per workgroup: 64 workgroups over 256
    tiles) captured in the substep graph and replayed give the eager
    one-tile-per-workgroup results bit for bit."""
    fams = pbf_amd._lib.FAMILIES
    scene = pbf_amd.config("dam16k", seed=3)
    outs = []
    for tiles, graph in ((1, False), (4, True)):
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_state(scene.x, scene.v)
        sim.set_friends_mode("direct")
        sim.set_workgroup_sizes({f: 64 for f in fams})
        sim.set_tiles_per_workgroup({f: tiles for f in fams})
        sim.set_graph(graph)
        sim.run(6)
        outs.append(sim.get_state())
        sim.close()
    for k in ("x", "v", "xs", "omega"):
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), k


def test_persistent_grid_strides_at_1m():
    """At 1M particles and 64-thread workgroups the persistent grid is
    smaller than the 16,384 tiles, so every workgroup walks several tiles of
    its XCD's run; results equal one workgroup per tile bit for bit."""
    fams = pbf_amd._lib.FAMILIES
    scene = pbf_amd.config("dam1m", seed=0)
    outs = []
    for t in (1, 0):
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_state(scene.x, scene.v)
        sim.set_friends_mode("direct")
        sim.set_workgroup_sizes({f: 64 for f in fams})
        sim.set_tiles_per_workgroup({f: t for f in fams})
        sim.run(2)
        outs.append(sim.get_state())
        sim.close()
    for k in ("x", "v", "xs", "omega"):
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), k


def test_autotune_restores_state_and_picks_sizes():
    scene = pbf_amd.config("dam16k", seed=8)
    ref_sim = pbf_amd.Simulation(scene.params, scene.options)
    ref_sim.set_state(scene.x, scene.v)
    ref_sim.run(3)
    want = ref_sim.get_state()
    ref_sim.close()
    sim = pbf_amd.Simulation(scene.params, scene.options)
    sim.set_state(scene.x, scene.v)
    chosen = sim.autotune(steps=2)
    assert set(chosen) == set(pbf_amd._lib.FAMILIES)
    assert all(v in (64, 128, 256, 512) for v in chosen.values())
    assert sim.stats()["workgroup_size"] == chosen
    assert all(t in (0, 1, 4) for t in sim.stats()["tiles_per_workgroup"].values())
    sim.run(3)
    got = sim.get_state()
    for k in ("x", "v", "xs"):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    sim.close()


def test_sparse_tiles_fall_back_to_global_scan():
    """Tiles whose neighbourhood does not fit in LDS scan global memory and
    give the same friends lists."""
    rng = np.random.default_rng(3)
    n = 6000
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.random((n, 3)).astype(np.float32) * np.float32(6.0)
    p = pbf_amd.physical_params(n, 0.1, (0, 0, 0), (6.0, 6.0, 6.0))
    scene = pbf_amd.Scene(p, pbf_amd.default_options(), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    sim.set_friends_mode("staged")
    advance(sim, ref, 2)
    assert sim.stats()["unstaged_tiles"] > 0
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


@pytest.mark.parametrize("mode", ["direct", "staged"])
@pytest.mark.parametrize("size", [64, 128, 256, 512])
def test_per_particle_friends_builds_match(mode, size):
    """Both friends builds (global reads, LDS-staged) give the oracle's
    lists at every tile size."""
    scene = pbf_amd.config("dam16k", seed=14)
    sim, ref = make_pair(scene)
    sim.set_friends_mode(mode)
    assert sim.stats()["friends_staged"] == sim.FRIENDS_MODES[mode]
    sim.set_workgroup_sizes({f: size for f in pbf_amd._lib.FAMILIES})
    advance(sim, ref, 3)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_friends_lanes_match(lanes):
    """The global-read friends build with 1 to 16 lanes per particle gives
    the oracle's lists: on the 16,384-particle dam break, on hash-folded
    buckets (gridBufSize 97, visited once per alias) and on a cluster that
    overflows the 50-per-circle cap (the first 50 in scan order are kept
    across the lanes' shares, every further friend is counted as dropped)."""
    scene = pbf_amd.config("dam16k", seed=21)
    sim, ref = make_pair(scene)
    sim.set_friends_lanes(lanes)
    assert sim.stats()["friends_lanes"] == lanes
    advance(sim, ref, 3)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()

    scene = pbf_amd.dam_break(12, 12, 12, 0.1, (1.2, 1.2, 1.2), seed=11, grid_buf_size=97)
    sim, ref = make_pair(scene)
    sim.set_friends_lanes(lanes)
    advance(sim, ref, 2)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()

    rng = np.random.default_rng(0)
    n = 400
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = (0.55 + 0.02 * rng.random((n, 3))).astype(np.float32)
    p = pbf_amd.physical_params(n, 0.1, (0, 0, 0), (1.0, 1.0, 1.0), iters=1)
    scene = pbf_amd.Scene(p, pbf_amd.default_options(), x, np.zeros_like(x))
    sim, ref = make_pair(scene)
    sim.set_friends_lanes(lanes)
    advance(sim, ref, 1)
    st = check_structure(sim, ref)
    flags, dropped = sim.error_flags(reset=True)
    counts = st["friends"][: 5 * n].reshape(5, n)
    assert flags & pbf_amd.PBF_FLAG_FRIENDS_OVERFLOW
    assert dropped == int(n * (n - 1) - counts.sum())
    sim.close()


def test_friends_lanes_default_and_validation():
    """Without a setting the build takes 16 lanes per particle up to two
    64-particle waves per CU (16 at 16,384 particles on the 256-CU device), 1
    above (1M); other counts than 0, 1, 2, 4, 8, 16 are refused."""
    for name, want in (("dam16k", 16), ("dam1m", 1)):
        scene = pbf_amd.config(name, seed=0)
        sim = pbf_amd.Simulation(scene.params, scene.options)
        sim.set_state(scene.x, scene.v)
        assert sim.stats()["friends_lanes"] == want, name
        sim.set_friends_mode("staged")   # the LDS-staged build runs one lane per particle
        assert sim.stats()["friends_lanes"] == 1
        with pytest.raises(pbf_amd.PbfError):
            sim.set_friends_lanes(3)
        sim.close()


def test_fast_sqrt_and_reciprocal_are_ieee():
    # every normal binary32 input: the kernels' square root and reciprocal
    # equal sqrtf and 1.0f / x
    import ctypes
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    fa, fb = ctypes.c_uint32(), ctypes.c_uint32()
    assert pbf_amd.lib().pbf_check_fast_math(ctypes.byref(a), ctypes.byref(b), ctypes.byref(fa), ctypes.byref(fb)) == 0
    assert (a.value, b.value) == (0, 0), (a.value, hex(fa.value), b.value, hex(fb.value))


@pytest.mark.parametrize("mode", ["direct", "staged"])
def test_friends_modes_agree_at_1m(mode):
    """Both friends build variants give the oracle's lists on the
    1,048,576-particle dam break after a step (the lists, not just the state:
    the sweeps only read the processed prefix)."""
    scene = pbf_amd.config("dam1m", seed=1)
    sim, ref = make_pair(scene)
    sim.set_friends_mode(mode)
    advance(sim, ref, 1)
    check_structure(sim, ref)
    check_state(sim, ref, scene)
    sim.close()
