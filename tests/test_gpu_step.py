"""GPU parity of the plain step, binning, sort order and proximity graph (engine vs oracle).

All engine calls go through the C ABI (cellflow_amd/libcellflow.so).  Decisions (neighbour
counts, cells, sort order, edges) must match the oracle bit for bit; FP32 positions,
velocities and accelerations must be within the tolerances stated in tests/helpers.py.
"""

import numpy as np
import pytest

from cellflow_amd import CellFlowError, ParticleSimulation, configs, particles_from_state
from helpers import check_step, check_step_edges, make_sim, small_cfg, state_of

pytestmark = pytest.mark.gpu

CASES = {
    # name: config overrides
    "uniform_3k": dict(n=3000, side=0.5),
    "dense_2k": dict(n=2000, side=0.25, ratio=0.6),
    "small_grid_generic": dict(n=800, side=0.12),        # 2 cells per axis: per-pair minimum image
    "tiny_grid_one_cell": dict(n=300, side=0.06),         # 1 cell per axis
    "ten_types": dict(n=3000, side=0.5, T=10),
    "one_type": dict(n=1500, side=0.4, T=1),
    "balance_over_one": dict(n=3000, side=0.45, balance=1.2, ratio=0.5),
    "negative_ratio": dict(n=3000, side=0.5, ratio=-1.5),  # some effR <= 0: no interaction
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_single_step_matches_oracle(orc, name):
    cfg = small_cfg(**CASES[name])
    st, phys, params = state_of(cfg)
    # non-zero previous counts and velocities exercise the density factor and friction
    rng = np.random.default_rng(7)
    st["prev"] = rng.integers(0, 600, len(st["type"])).astype(np.int32)
    st["vel"] = rng.normal(0, 0.3, st["pos"].shape).astype(np.float32)
    sim = make_sim(st, phys)
    sim.simulate(params, 1)
    got = sim.getParticleData()
    counts = sim.getNeighborCounts()
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"], mode=0)
    assert np.array_equal(counts, ref["count"])
    assert np.array_equal(got["ptype"], st["type"])
    check_step(got, ref, phys, st["vel"])
    sim.close()


def test_non_cubic_canvas(orc):
    cfg = small_cfg(n=2500)
    cfg["canvas"] = [0.7, 0.3, 0.45]
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.simulate(params, 1)
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"], mode=0)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    check_step(sim.getParticleData(), ref, phys, st["vel"])


@pytest.mark.parametrize("sort", [1, 0])
def test_sort_order_matches_oracle(orc, sort):
    cfg = small_cfg(n=5000, side=0.6)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.setTuning(sort_variant=sort)
    sim.simulate(params, 3)
    dims, reach = sim.getGrid()
    pos, _, _ = sim.getState()
    cells, index = sim.getSortOrder()
    ref_cells = orc.bin_cells(pos, dims, phys["canvas"])
    order = orc.order(ref_cells)
    assert np.array_equal(index, order)
    assert np.array_equal(cells, ref_cells[order])
    # every interacting pair lies within `reach` cells (periodic) of each other
    assert reach == 1


def assert_sorted_like_oracle(orc, sim, phys):
    dims, _ = sim.getGrid()
    pos, _, _ = sim.getState()
    cells, index = sim.getSortOrder()
    ref_cells = orc.bin_cells(pos, dims, phys["canvas"])
    order = orc.order(ref_cells)
    assert np.array_equal(index, order)
    assert np.array_equal(cells, ref_cells[order])


def step_and_check(orc, st, phys, params, sort, steps=3):
    """One step against the oracle, then the canonical order after each of `steps` steps."""
    sim = make_sim(st, phys)
    sim.setTuning(sort_variant=sort)
    sim.simulate(params, 1)
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"], mode=0)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    check_step(sim.getParticleData(), ref, phys, st["vel"])
    assert_sorted_like_oracle(orc, sim, phys)
    for _ in range(steps - 1):
        sim.simulate(params, 1)
        assert_sorted_like_oracle(orc, sim, phys)
    sim.close()


@pytest.mark.parametrize("sort", [1, 0])
def test_far_movers(orc, sort):
    """Velocities of several cells per step: most particles leave their tile's window and
    are placed through the far list."""
    cfg = small_cfg(n=3000, side=0.5)
    st, phys, params = state_of(cfg)
    rng = np.random.default_rng(11)
    st["vel"] = rng.normal(0, 300.0, st["pos"].shape).astype(np.float32)  # ~0.3 per step
    step_and_check(orc, st, phys, params, sort)


@pytest.mark.parametrize("sort", [1, 0])
def test_one_cell_beyond_the_lds(orc, sort):
    """3000 particles in one cell: more than a tile sorts in LDS (the one-cell fallback)."""
    cfg = small_cfg(n=3000, side=0.5)
    st, phys, params = state_of(cfg)
    rng = np.random.default_rng(12)
    st["pos"] = (0.1 + rng.random(st["pos"].shape) * 0.02).astype(np.float32)
    step_and_check(orc, st, phys, params, sort, steps=2)


@pytest.mark.parametrize("sort", [1, 0])
def test_tile_beyond_the_lds(orc, sort):
    """6000 particles in one row of cells: one tile holds more than the LDS takes, no single
    cell does (the tile is sorted a few cells at a time)."""
    cfg = small_cfg(n=6000, side=0.5)
    st, phys, params = state_of(cfg)
    rng = np.random.default_rng(13)
    pos = st["pos"].copy()
    pos[:, 0] = rng.random(len(pos)) * 0.5
    pos[:, 1] = 0.2 + rng.random(len(pos)) * 0.01
    pos[:, 2] = 0.3 + rng.random(len(pos)) * 0.01
    st["pos"] = pos.astype(np.float32)
    step_and_check(orc, st, phys, params, sort)


def test_teacher_forced_steps(orc):
    """Each of 6 steps is compared with an oracle step from the engine's previous state."""
    cfg = small_cfg(n=4000, side=0.5)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    prev = st["prev"]
    for k in range(6):
        before = sim.getParticleData()
        prev = sim.getNeighborCounts()
        sim.simulate(params, 1)
        got = sim.getParticleData()
        ref = orc.step(before["pos"], before["vel"], before["ptype"], prev, phys, st["table"],
                       st["rbt"], mode=0)
        assert np.array_equal(sim.getNeighborCounts(), ref["count"]), f"step {k}"
        check_step(got, ref, phys, before["vel"])


def test_c2_first_step_and_edges(orc):
    cfg = configs.load("c2_100k_edges")
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    e = cfg["edges"]
    sim.simulate(params, 1, edges=(e["distance"], e["max_connections"]))
    got = sim.getParticleData()
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"], mode=1)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    check_step(got, ref, phys, st["vel"])
    edges = sim.getEdges()
    ref_edges = orc.edges(got["pos"], got["ptype"], e["distance"], e["max_connections"], mode=1)
    assert len(edges) > 1000
    assert np.array_equal(edges, ref_edges)


def clustered_state(n=6000, side=1.0, T=2, blobs=6, sd=0.01, seed=3):
    cfg = small_cfg(n=n, side=side, T=T, seed=seed)
    cfg["init"] = dict(kind="blobs", count=blobs, sd=sd)
    return state_of(cfg)


@pytest.mark.parametrize("maxc", [1, 4, 16, 20, 32])
def test_edges_truncation_matches_oracle(orc, maxc):
    """Dense blobs: most owners have far more than 2*maxc qualifying partners, so the
    index-order truncation and the distance sort both matter."""
    st, phys, params = clustered_state()
    sim = make_sim(st, phys)
    sim.simulate(params, 0)  # sets the grid from the parameters without stepping
    for dist in (0.004, 0.02, 0.15):  # 0.15 exceeds the cell size: wider edge stencil
        edges = sim.generateProximityGraph(dist, maxc)
        ref = orc.edges(st["pos"], st["type"], dist, maxc, mode=0)
        assert np.array_equal(edges, ref), f"dist={dist} maxc={maxc}"


@pytest.mark.parametrize("variant,lanes,divisor", [(0, 4, 1), (0, 16, 1), (3, 8, 1), (3, 16, 1),
                                                   (3, 32, 1), (4, 8, 1), (3, 8, 2), (4, 8, 2)])
def test_fused_edges_dense_blobs(orc, variant, lanes, divisor):
    """Fused edge candidates on dense blobs: many owners exceed the per-particle LDS edge
    list (variants 3 and 4 at 8 lanes then select in the wave's LDS lists, with the 27- and
    the 125-cell stencil; the others take the streaming selection) and, in variants 3 and 4,
    hold several candidate-list chunks; the graph must still equal the oracle's."""
    st, phys, params = clustered_state(n=5000, blobs=4, sd=0.006)
    sim, twin = make_sim(st, phys), make_sim(st, phys)
    for sort in (1, 0):
        for h in (sim, twin):
            h.setTuning(cell_divisor=divisor, force_variant=variant, lanes_per_particle=lanes,
                        sort_variant=sort)
        check_step_edges(orc, sim, twin, params, (0.01, 16), steps=2)
        twin.simulate(params, 1)  # back in step with sim


@pytest.mark.parametrize("variant", [3, 4])
def test_fused_edges_past_the_wave_lists(orc, variant):
    """One blob so dense that most owners have more qualifying partners than the wave's LDS
    lists hold (320): those take the exact streaming selection, the few with fewer the
    wave's; the graph must equal the oracle's."""
    st, phys, params = clustered_state(n=4000, blobs=1, sd=0.002)
    sim, twin = make_sim(st, phys), make_sim(st, phys)
    for h in (sim, twin):
        h.setTuning(force_variant=variant, lanes_per_particle=8)
    assert check_step_edges(orc, sim, twin, params, (0.01, 16), steps=2) > 0


def test_fused_edges_every_step(orc):
    """With edges on, each step's graph is of the post-step state (the last one is kept by
    the standalone pass; the one before it was built inside the last step)."""
    cfg = small_cfg(n=4000, side=0.5)
    st, phys, params = state_of(cfg)
    sim, twin = make_sim(st, phys), make_sim(st, phys)
    for k in range(3):
        check_step_edges(orc, sim, twin, params, (0.03, 5), steps=2 + k)
        twin.simulate(params, 1)  # back in step with sim


def test_step_edges_state_errors():
    """The in-step graph exists only after a call of two or more steps with edges, and a
    newly loaded state drops it."""
    st, phys, params = state_of(small_cfg(n=2000, side=0.4))
    sim = make_sim(st, phys)
    for nsteps, edges in ((1, (0.03, 5)), (3, None)):
        sim.simulate(params, nsteps, edges=edges)
        with pytest.raises(CellFlowError):
            sim.getStepEdges()
    sim.simulate(params, 2, edges=(0.03, 5))
    assert len(sim.getStepEdges()) > 0
    sim.setParticles(sim.getParticleData())
    with pytest.raises(CellFlowError):
        sim.getStepEdges()


def test_run_twice_bit_identical():
    cfg = small_cfg(n=5000, side=0.55)
    st, phys, params = state_of(cfg)
    outs = []
    for _ in range(2):
        sim = make_sim(st, phys)
        sim.simulate(params, 25, edges=(0.02, 8))
        outs.append((sim.getParticleData().tobytes(), sim.getNeighborCounts().tobytes(),
                     sim.getEdges().tobytes()))
        sim.close()
    assert outs[0] == outs[1]


def test_resume_bit_identical():
    """20 steps == 10 steps + snapshot (records + counts) into a new handle + 10 steps."""
    cfg = small_cfg(n=4000, side=0.5)
    st, phys, params = state_of(cfg)
    a = make_sim(st, phys)
    a.simulate(params, 20)
    b = make_sim(st, phys)
    b.simulate(params, 10)
    snap, cnt = b.getParticleData(), b.getNeighborCounts()
    c = make_sim(st, phys)
    c.setParticles(snap)
    c.setNeighborCounts(cnt)
    c.simulate(params, 10)
    assert a.getParticleData().tobytes() == c.getParticleData().tobytes()
    assert np.array_equal(a.getNeighborCounts(), c.getNeighborCounts())


def test_line_vertices_layout():
    cfg = small_cfg(n=3000, side=0.4, T=3)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.03, 4))
    edges = sim.getEdges()
    parts = sim.getParticleData()
    colors = np.array([[0.9, 0.1, 0.2], [0.1, 0.8, 0.3], [0.2, 0.3, 0.7]], np.float32)
    v = sim.getLineVertices(colors)
    assert v.shape == (2 * len(edges), 6)
    i, j = edges[:, 0], edges[:, 1]
    assert np.array_equal(v[0::2, :3], parts["pos"][i])
    assert np.array_equal(v[1::2, :3], parts["pos"][j])
    col = colors[parts["ptype"][i] % 3]
    assert np.array_equal(v[0::2, 3:], col) and np.array_equal(v[1::2, 3:], col)


def test_seeded_init_matches_generator():
    n, T = 5000, 7
    sim = ParticleSimulation(n, T)
    canvas = (8000.0, 8000.0, 8000.0)
    sim.initializeParticles(11, canvas, 2000.0)
    p = sim.getParticleData()
    u = configs.uniform_f32(11, configs.S_POS, 3 * n).reshape(n, 3)
    spawn = np.float32(2000.0)
    off = (np.float32(8000.0) - spawn) * np.float32(0.5)
    assert np.array_equal(p["pos"], off + u * spawn)
    t = np.minimum((configs.uniform_f32(11, configs.S_TYPE, n) * np.float32(T)).astype(np.uint32), T - 1)
    assert np.array_equal(p["ptype"], t)
    assert not p["vel"].any() and not p["acc"].any()


def test_move_universe_matches_oracle(orc):
    """moveUniverse: positions shift with the reference's FP32 wrap (bit-exact), velocities,
    counts and accelerations stay, the binning is redone canonically and the next step
    still matches the oracle."""
    cfg = small_cfg(n=4000, side=0.5)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.simulate(params, 2)
    before = sim.getParticleData()
    counts = sim.getNeighborCounts()
    sim.moveUniverse(0.137, -0.41, 0.05, canvas=phys["canvas"])
    after = sim.getParticleData()
    want = orc.move(before["pos"], 0.137, -0.41, 0.05, phys["canvas"])
    assert np.array_equal(after["pos"], want)
    for f in ("vel", "acc", "ptype", "pad"):
        assert np.array_equal(after[f], before[f]), f
    assert np.array_equal(sim.getNeighborCounts(), counts)
    dims, _ = sim.getGrid()
    cells, index = sim.getSortOrder()
    ref_cells = orc.bin_cells(want, dims, phys["canvas"])
    assert np.array_equal(index, orc.order(ref_cells))
    ref = orc.step(want, after["vel"], after["ptype"], counts, phys, st["table"], st["rbt"],
                   mode=0)
    sim.simulate(params, 1)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    check_step(sim.getParticleData(), ref, phys, after["vel"])
    # the reference's default: dz = 0 on the default 8000^3 canvas
    big = ParticleSimulation(1000, 3)
    big.initializeParticles(5)
    p0 = big.getParticleData()["pos"]
    big.moveUniverse(-2500.0, 7000.0)
    assert np.array_equal(big.getParticleData()["pos"], orc.move(p0, -2500.0, 7000.0, 0.0,
                                                                 [8000.0] * 3))


def test_nonfinite_state_reported():
    cfg = small_cfg(n=500, side=0.5)
    st, phys, params = state_of(cfg)
    st["vel"][7] = [np.inf, 0, 0]
    sim = make_sim(st, phys)
    with pytest.raises(CellFlowError) as ei:
        sim.simulate(params, 1)
    assert ei.value.code == 4


def test_edge_capacity_error():
    cfg = small_cfg(n=2000, side=0.3)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.05, 8))
    import ctypes as C
    from cellflow_amd import _abi
    cnt = C.c_longlong()
    buf = np.zeros((1, 2), np.int32)
    rc = _abi.lib().cf_get_edges(sim._h, buf.ctypes.data_as(C.POINTER(C.c_int32)), 1, C.byref(cnt))
    assert rc == _abi.CF_ERR_CAPACITY and cnt.value > 1


# every combination the autotuner times: (variant, lanes, divisor, block size)
VARIANTS = [(0, g, d, 128) for d in (1, 2) for g in (4, 8, 16, 32)] + \
    [(1, 16, d, b) for d in (1, 2) for b in (64, 128, 256)] + [(2, 16, 1, 128), (2, 16, 2, 128)] + \
    [(3, g, 1, 128) for g in (4, 8, 16, 32)] + [(3, g, 2, 128) for g in (8, 16, 32)] + \
    [(4, g, d, 128) for d in (1, 2) for g in (8, 16)]


@pytest.mark.parametrize("sort", [1, 0])
@pytest.mark.parametrize("variant,lanes,divisor,block", VARIANTS)
@pytest.mark.parametrize("name", ["uniform_3k", "small_grid_generic", "dense_2k"])
def test_variants_match_oracle(orc, variant, lanes, divisor, block, name, sort):
    """Every force-kernel variant, lane grouping and stencil reach, with either step sort,
    gives the oracle's counts, forces and edges."""
    cfg = small_cfg(**CASES[name])
    st, phys, params = state_of(cfg)
    rng = np.random.default_rng(11)
    st["prev"] = rng.integers(0, 600, len(st["type"])).astype(np.int32)
    sim = make_sim(st, phys)
    sim.setTuning(cell_divisor=divisor, force_variant=variant, lanes_per_particle=lanes,
                  block_size=block, sort_variant=sort)
    sim.simulate(params, 1, edges=(0.02, 6))
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"], mode=0)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    got = sim.getParticleData()
    check_step(got, ref, phys, st["vel"])
    assert np.array_equal(sim.getEdges(), orc.edges(got["pos"], got["ptype"], 0.02, 6, mode=0))
    # two more steps: the second builds the graph of the state between them inside its
    # force pass and sort (the fused path); a twin handle at the same state supplies it
    twin = make_sim(st, phys)
    twin.setTuning(cell_divisor=divisor, force_variant=variant, lanes_per_particle=lanes,
                   block_size=block, sort_variant=sort)
    twin.simulate(params, 1, edges=(0.02, 6))
    check_step_edges(orc, sim, twin, params, (0.02, 6))


def test_consolidated_tail_bit_identical_to_list_kernel():
    """Variant 4 differs from variant 3 only in which lanes integrate: the rollouts (state,
    counts, both graphs) are bit for bit the same."""
    st, phys, params = state_of(small_cfg(n=5000, side=0.6))
    out = []
    for variant in (3, 4):
        sim = make_sim(st, phys)
        sim.setTuning(force_variant=variant, lanes_per_particle=8)
        sim.simulate(params, 12, edges=(0.03, 8))
        pos, vel, ids = sim.getState()
        out.append((pos.tobytes(), vel.tobytes(), ids.tobytes(), sim.getEdges().tobytes(),
                    sim.getStepEdges().tobytes(), sim.getNeighborCounts().tobytes(),
                    sim.getParticleData().tobytes()))
        sim.close()
    assert out[0] == out[1]


@pytest.mark.parametrize("variant,lanes", [(0, 16), (3, 8), (4, 8)])
def test_both_sorts_bit_identical(variant, lanes):
    """The two step sorts produce the same canonical order, so a rollout is bit for bit the
    same with either (here with fused edges every step)."""
    st, phys, params = state_of(small_cfg(n=5000, side=0.6))
    out = []
    for sort in (1, 0):
        sim = make_sim(st, phys)
        sim.setTuning(force_variant=variant, lanes_per_particle=lanes, sort_variant=sort)
        sim.simulate(params, 12, edges=(0.03, 8))
        pos, vel, ids = sim.getState()
        out.append((pos.tobytes(), vel.tobytes(), ids.tobytes(), sim.getEdges().tobytes(),
                    sim.getNeighborCounts().tobytes()))
        sim.close()
    assert out[0] == out[1]


def test_profile_steps_times_each_kernel_and_steps_the_state():
    """profileSteps times the step's kernels with HIP events and advances the state exactly
    as simulate does (same kernels, launched one by one instead of from a graph)."""
    st, phys, params = state_of(small_cfg(n=4000, side=0.5))
    a, b = make_sim(st, phys), make_sim(st, phys)
    t = a.profileSteps(params, 6, edges=(0.02, 8))
    assert list(t) == ["force", "sort", "edges", "edge_emit"]
    assert all(v > 0.0 for v in t.values())
    b.simulate(params, 6, edges=(0.02, 8))
    assert a.getParticleData().tobytes() == b.getParticleData().tobytes()
    assert np.array_equal(a.getEdges(), b.getEdges())
    c = make_sim(st, phys)
    c.setTuning(sort_variant=0)
    t0 = c.profileSteps(params, 3)
    assert list(t0) == ["force", "scan", "fill", "rank_scatter"]
    assert all(v > 0.0 for v in t0.values())


def test_gets_into_caller_buffers():
    """getParticleData / getLineVertices / getTriangleVertices fill a caller's array in place
    (as the reference fills the caller's vector) and give the same bytes as a fresh array; a
    buffer that cannot be filled in place (too small, not contiguous, wrong dtype) raises
    ValueError and is left untouched."""
    cfg = small_cfg(n=3000, side=0.4, T=3)
    st, phys, params = state_of(cfg)
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.03, 4))
    colors = np.eye(3, dtype=np.float32)
    parts, lines = sim.getParticleData(), sim.getLineVertices(colors)
    sim.generateTriangleMesh(0.03, 6)
    tverts = sim.getTriangleVertices(colors)
    assert len(tverts) > 0
    pbuf = np.zeros(len(parts) + 7, parts.dtype)
    vbuf = np.zeros(lines.size + 6, np.float32).reshape(-1, 6)  # 2-D and contiguous: usable
    tbuf = np.zeros(tverts.size + 9, np.float32)
    p2, l2 = sim.getParticleData(out=pbuf), sim.getLineVertices(colors, out=vbuf)
    t2 = sim.getTriangleVertices(colors, out=tbuf)
    assert np.shares_memory(p2, pbuf) and np.shares_memory(l2, vbuf)
    assert np.shares_memory(t2, tbuf)
    assert p2.tobytes() == parts.tobytes() and np.array_equal(l2, lines)
    assert np.array_equal(t2, tverts)
    bad = [np.zeros(10, parts.dtype),                         # too small
           np.zeros(2 * (len(parts) + 1), parts.dtype)[::2],   # not contiguous
           np.zeros(len(parts) + 1, np.float32)]               # wrong dtype
    for buf in bad:
        with pytest.raises(ValueError):
            sim.getParticleData(out=buf)
        assert not buf.tobytes().strip(b"\0")
    strided = np.zeros((lines.size + 6) * 2, np.float32)[::2]
    with pytest.raises(ValueError):
        sim.getLineVertices(colors, out=strided)
    assert not strided.any()
    with pytest.raises(ValueError):
        sim.getTriangleVertices(colors, out=np.zeros(tverts.size - 1, np.float32))
