"""GPU: the batched ensemble and HIP-graph replays (engine vs engine, bit for bit).

An ensemble runs every simulation through the same kernels in the same per-particle order
as a stand-alone plain run, so each member must equal its plain run bit for bit.  Captured
step graphs must equal direct launches bit for bit on every replay.
"""

import numpy as np
import pytest

from cellflow_amd import Ensemble, SimulationParams, configs
from helpers import make_sim

pytestmark = pytest.mark.gpu


def small_ensemble(S=6, seed=3):
    cfg = dict(configs.load("c4_ensemble_256"), n_min=300, n_max=4000)
    return configs.make_ensemble(cfg, seed=seed, sims=S)


@pytest.mark.parametrize("variant,lanes,divisor", [(0, 16, 1), (0, 8, 2), (3, 4, 1), (4, 8, 1)])
def test_ensemble_members_equal_plain_runs(variant, lanes, divisor):
    """Also for the list kernels, whose blocks share one pair table (a block's particles
    always lie in one simulation)."""
    sims = small_ensemble()
    ens = Ensemble(sims)
    ens.setTuning(cell_divisor=divisor, force_variant=variant, lanes_per_particle=lanes)
    ens.simulate(12)
    for s, (ph, st) in enumerate(sims):
        plain = make_sim(st, ph)
        plain.setTuning(cell_divisor=divisor, force_variant=variant, lanes_per_particle=lanes)
        plain.simulate(SimulationParams.from_physics(ph), 12)
        a, b = ens.getParticleData(s), plain.getParticleData()
        assert a.tobytes() == b.tobytes(), f"sim {s}"
        assert np.array_equal(ens.getNeighborCounts(s), plain.getNeighborCounts())
        plain.close()
    ens.close()


def test_ensemble_matches_oracle_first_step(orc):
    from helpers import check_step
    sims = small_ensemble(S=4, seed=8)
    ens = Ensemble(sims)
    ens.simulate(1)
    for s, (ph, st) in enumerate(sims):
        ref = orc.step(st["pos"], st["vel"], st["type"], st["prev"], ph, st["table"], st["rbt"],
                       mode=0)
        assert np.array_equal(ens.getNeighborCounts(s), ref["count"])
        check_step(ens.getParticleData(s), ref, ph, st["vel"])


@pytest.mark.parametrize("edges", [None, (0.02, 8)])
def test_graph_replays_equal_direct_launches(edges):
    """Replay a captured 4-step graph several times with eager work (state reads,
    synchronisation) in between; every replay must equal direct launches."""
    from helpers import small_cfg, state_of
    st, phys, params = state_of(small_cfg(n=4000, side=0.5))
    direct = make_sim(st, phys)
    direct.setTuning(steps_per_graph=1)
    graph = make_sim(st, phys)
    graph.setTuning(steps_per_graph=4)
    for rep in range(4):
        direct.simulate(params, 9, edges=edges)
        graph.simulate(params, 9, edges=edges)
        assert direct.getParticleData().tobytes() == graph.getParticleData().tobytes(), rep
        assert np.array_equal(direct.getNeighborCounts(), graph.getNeighborCounts())
        if edges:
            assert np.array_equal(direct.getEdges(), graph.getEdges())
