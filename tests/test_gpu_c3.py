"""GPU: the 1,000,000-particle clustered workload at full size
(cellflow_amd/configs/c3_1m_clustered.json: 200 Gaussian blobs of sd 0.05 in a periodic cube
of side 6, radius 0.04, ratio 0.5, balance 1.2, edges 0.02/16).  Cells hold hundreds of
particles and each particle has ~200 neighbours, the regime where the pair loop dominates.

The first step is checked against the oracle (counts, cells and edges bit-exact, FP32 state
within tolerance); the following steps by properties that hold at any size: the neighbour
relation is symmetric (even count sum), the binning is canonical, the edge list is
canonical and obeys the selection rule, and a teacher-forced step from the engine's own
state still matches the oracle."""

import numpy as np
import pytest

from cellflow_amd import SimulationParams, configs
from helpers import check_step, make_sim

pytestmark = pytest.mark.gpu


def test_c3_full_size(orc):
    cfg = configs.load("c3_1m_clustered")
    st, phys = configs.make_state(cfg), configs.physics(cfg)
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    params = SimulationParams.from_physics(phys)
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=edges)
    ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], phys, st["table"], st["rbt"],
                   mode=1)
    counts = sim.getNeighborCounts()
    assert np.array_equal(counts, ref["count"])
    got = sim.getParticleData()
    err = check_step(got, ref, phys, st["vel"])
    ref_edges = orc.edges(got["pos"], got["ptype"], *edges, mode=1)
    assert np.array_equal(sim.getEdges(), ref_edges)
    dims, _ = sim.getGrid()
    cells, index = sim.getSortOrder()
    occ = np.bincount(cells, minlength=int(np.prod(dims)))
    print(f"step 1: mean count {counts.mean():.1f} max {counts.max()}, max cell occupancy "
          f"{occ.max()}, edges {len(ref_edges)}, velocity error {err:.3g} of scale")

    sim.simulate(params, 19, edges=edges)
    state = sim.getParticleData()
    prev = sim.getNeighborCounts()
    assert prev.sum() % 2 == 0
    dims, _ = sim.getGrid()
    cells, index = sim.getSortOrder()
    ref_cells = orc.bin_cells(state["pos"], dims, phys["canvas"])
    assert np.array_equal(index, orc.order(ref_cells))
    E = sim.getEdges()
    assert np.array_equal(E, orc.edges(state["pos"], state["ptype"], *edges, mode=1))
    # teacher-forced step 21 from the engine's own state
    ref = orc.step(state["pos"], state["vel"], state["ptype"], prev, phys, st["table"],
                   st["rbt"], mode=1)
    sim.simulate(params, 1)
    assert np.array_equal(sim.getNeighborCounts(), ref["count"])
    err = check_step(sim.getParticleData(), ref, phys, state["vel"])
    print(f"step 21: mean count {ref['count'].mean():.1f}, edges {len(E)}, velocity error "
          f"{err:.3g} of scale")
