"""GPU parity of the five workloads at their stated sizes (cellflow_amd/configs/*.json).

* 10k plain (10,000 particles, 6 types, cube of side 1): the first step against the oracle,
  then its whole 100-step run with a teacher-forced step 101 from the engine's own state.
* 100k with edges (100,000 particles, edges 0.02 / 16): all 1,000 steps, then a
  teacher-forced step 1,001; the graphs of the states after steps 999 and 1,000 (the one
  built inside step 1,000's force pass and sort, and the standalone one) against the
  oracle's; binning and order canonical.  Run with the tuning the autotuner picks for this
  workload (list kernel with the consolidated tail, 8 lanes per particle, tile sort) and
  with the default kernel.
* 1M clustered (200 Gaussian blobs, edges): all 500 steps, then the same checks.
* 256-simulation ensemble (2k-50k particles each): every member's first step against the
  oracle, then 10 more steps with members equal to their plain runs bit for bit.

Decisions (counts, cells, order, edges) must be bit-exact; FP32 state within the one-step
tolerances of tests/helpers.py.  A teacher-forced step starts the oracle from the engine's
FP32 state, so it checks the step at the evolved state, where the trajectory itself can no
longer be compared pointwise (the dynamics are chaotic: see test_gpu_horizon.py).
"""

import numpy as np
import pytest

from cellflow_amd import Ensemble, SimulationParams, configs
from helpers import check_step, make_sim

pytestmark = pytest.mark.gpu

# The autotuner's pick for the 100k workload with edges (the list kernel with its
# integration tail consolidated per block), and the engine's default tuning.
LIST_TILE = dict(cell_divisor=1, force_variant=4, lanes_per_particle=8, sort_variant=1)
DEFAULT = dict(cell_divisor=1, force_variant=0, lanes_per_particle=16, sort_variant=1)


def workload(name):
    cfg = configs.load(name)
    st, phys = configs.make_state(cfg), configs.physics(cfg)
    return cfg, st, phys, SimulationParams.from_physics(phys)


def assert_canonical(orc, sim, phys, state):
    dims, _ = sim.getGrid()
    cells, index = sim.getSortOrder()
    ref_cells = orc.bin_cells(state["pos"], dims, phys["canvas"])
    order = orc.order(ref_cells)
    assert np.array_equal(index, order)
    assert np.array_equal(cells, ref_cells[order])
    return np.bincount(cells, minlength=int(np.prod(dims)))


def teacher_forced(orc, sim, phys, table, rbt, params, mode=1):
    """One engine step from its current state against an oracle step from the same FP32
    state; returns the velocity error as a fraction of its scale."""
    before = sim.getParticleData()
    prev = sim.getNeighborCounts()
    assert prev.sum() % 2 == 0  # the neighbour relation is symmetric
    ref = orc.step(before["pos"], before["vel"], before["ptype"], prev, phys, table, rbt,
                   mode=mode)
    sim.simulate(params, 1)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    return check_step(sim.getParticleData(), ref, phys, before["vel"])


def edge_candidates(pos, ptype, dist):
    """Per owner, the number of same-type particles of larger index within `dist`
    (non-periodic, FP32 squared distance as the engine computes it): the qualifying
    candidates before truncation (ParticleSimulation.cu:214-227)."""
    from scipy.spatial import cKDTree
    pos = np.asarray(pos, np.float32)
    d2max = np.float32(dist) * np.float32(dist)
    cnt = np.zeros(len(pos), np.int64)
    for t in np.unique(ptype):
        idx = np.nonzero(ptype == t)[0]
        pairs = cKDTree(pos[idx].astype(np.float64)).query_pairs(float(dist) * 1.001,
                                                                output_type="ndarray")
        a, b = idx[pairs[:, 0]], idx[pairs[:, 1]]
        d = pos[a] - pos[b]
        s2 = d[:, 2] * d[:, 2] + (d[:, 1] * d[:, 1] + d[:, 0] * d[:, 0])
        keep = s2 < d2max
        np.add.at(cnt, np.minimum(a, b)[keep], 1)
    return cnt


def test_c1_full_workload(orc):
    cfg, st, phys, params = workload("c1_plain_10k")
    assert cfg["n"] == 10000 and cfg["canvas"] == [1.0, 1.0, 1.0]
    sim = make_sim(st, phys)
    sim.simulate(params, 1)
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"],
                   mode=1)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    err1 = check_step(sim.getParticleData(), ref, phys, st["vel"])
    sim.simulate(params, cfg["steps"] - 1)
    assert_canonical(orc, sim, phys, sim.getParticleData())
    err = teacher_forced(orc, sim, phys, st["table"], st["rbt"], params)
    print(f"10k: step 1 velocity error {err1:.3g}, step {cfg['steps'] + 1} {err:.3g} of scale; "
          f"mean count {sim.getNeighborCounts().mean():.2f}")
    sim.close()


@pytest.mark.parametrize("tuning", [LIST_TILE, DEFAULT], ids=["list_tile", "default"])
def test_c2_after_all_steps(orc, tuning):
    cfg, st, phys, params = workload("c2_100k_edges")
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    steps = cfg["steps"]
    sim, twin = make_sim(st, phys), make_sim(st, phys)
    for h in (sim, twin):
        h.setTuning(**tuning)
    sim.simulate(params, steps, edges=edges)       # last step builds the graph of S_999
    twin.simulate(params, steps - 1, edges=edges)  # S_999 and its standalone graph
    mid = twin.getParticleData()
    ref_mid = orc.edges(mid["pos"], mid["ptype"], *edges, mode=1)
    assert np.array_equal(twin.getEdges(), ref_mid)
    assert np.array_equal(sim.getStepEdges(), ref_mid)
    end = sim.getParticleData()
    ref_end = orc.edges(end["pos"], end["ptype"], *edges, mode=1)
    assert np.array_equal(sim.getEdges(), ref_end)
    occ = assert_canonical(orc, sim, phys, end)
    cand = edge_candidates(mid["pos"], mid["ptype"], edges[0])
    err = teacher_forced(orc, sim, phys, st["table"], st["rbt"], params)
    print(f"100k after {steps} steps ({tuning}): edges {len(ref_mid)} / {len(ref_end)}, "
          f"edge candidates per owner mean {cand.mean():.2f} max {cand.max()} "
          f"(over 24: {(cand > 24).sum()}), max cell occupancy {occ.max()}, mean count "
          f"{sim.getNeighborCounts().mean():.2f}, step {steps + 1} velocity error {err:.3g}")
    sim.close()
    twin.close()


def test_c3_after_all_steps(orc):
    cfg, st, phys, params = workload("c3_1m_clustered")
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    steps = cfg["steps"]
    sim, twin = make_sim(st, phys), make_sim(st, phys)
    for h in (sim, twin):  # the autotuner's pick for this workload
        h.setTuning(cell_divisor=1, force_variant=0, lanes_per_particle=16, sort_variant=0)
    sim.simulate(params, steps, edges=edges)
    twin.simulate(params, steps - 1, edges=edges)
    mid = twin.getParticleData()
    ref_mid = orc.edges(mid["pos"], mid["ptype"], *edges, mode=1)
    assert np.array_equal(twin.getEdges(), ref_mid)
    assert np.array_equal(sim.getStepEdges(), ref_mid)
    twin.close()
    end = sim.getParticleData()
    ref_end = orc.edges(end["pos"], end["ptype"], *edges, mode=1)
    assert np.array_equal(sim.getEdges(), ref_end)
    occ = assert_canonical(orc, sim, phys, end)
    err = teacher_forced(orc, sim, phys, st["table"], st["rbt"], params)
    counts = sim.getNeighborCounts()
    print(f"1M clustered after {steps} steps: edges {len(ref_end)}, max cell occupancy "
          f"{occ.max()}, mean count {counts.mean():.1f} max {counts.max()}, step {steps + 1} "
          f"velocity error {err:.3g} of scale")
    sim.close()


@pytest.mark.parametrize("tuning", [dict(force_variant=0, lanes_per_particle=4, sort_variant=0),
                                    dict(force_variant=0, lanes_per_particle=16, sort_variant=1)],
                         ids=["autotuned", "default"])
def test_c4_every_member(orc, tuning):
    cfg = configs.load("c4_ensemble_256")
    sims = configs.make_ensemble(cfg)
    assert len(sims) == 256
    sizes = np.array([len(s["type"]) for _, s in sims])
    assert sizes.min() >= cfg["n_min"] and sizes.max() <= cfg["n_max"]
    ens = Ensemble(sims)
    ens.setTuning(**tuning)
    ens.simulate(1)
    worst = 0.0
    for s, (ph, st) in enumerate(sims):
        ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], ph, st["table"], st["rbt"],
                       mode=1)
        assert np.array_equal(ens.getNeighborCounts(s), ref["count"]), f"member {s}"
        worst = max(worst, check_step(ens.getParticleData(s), ref, ph, st["vel"]))
    ens.simulate(10)
    # members of every size class against their stand-alone runs, bit for bit
    pick = sorted(set(np.argsort(sizes)[np.linspace(0, len(sizes) - 1, 10).astype(int)]))
    for s in pick:
        ph, st = sims[s]
        plain = make_sim(st, ph)
        plain.setTuning(**tuning)
        plain.simulate(SimulationParams.from_physics(ph), 11)
        assert ens.getParticleData(s).tobytes() == plain.getParticleData().tobytes(), s
        assert np.array_equal(ens.getNeighborCounts(s), plain.getNeighborCounts())
        plain.close()
    print(f"256-member ensemble ({sizes.sum()} particles, {sizes.min()}..{sizes.max()}): "
          f"every first step matches, worst velocity error {worst:.3g} of scale; "
          f"{len(pick)} members equal their plain runs after 11 steps")
    ens.close()
