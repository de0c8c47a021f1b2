"""GPU: the autotuner.  Tuning measures candidates on a copy of the state and must leave
the simulation exactly as it found it (bit for bit); afterwards the simulation must run
exactly like a fresh one given the chosen tuning directly.  The JSON cache answers the
second call for the same workload bucket."""

import json

import numpy as np
import pytest

from cellflow_amd import Ensemble, SimulationParams, configs
from helpers import make_sim, small_cfg, state_of

pytestmark = pytest.mark.gpu


def snapshot(sim):
    return sim.getParticleData().tobytes(), sim.getNeighborCounts().tobytes()


@pytest.mark.parametrize("edges", [None, (0.02, 8)])
def test_autotune_leaves_state_and_matches_direct_tuning(tmp_path, edges):
    st, phys, params = state_of(small_cfg(n=6000, side=0.6))
    sim = make_sim(st, phys)
    sim.simulate(params, 3, edges=edges)
    before = snapshot(sim)
    res = sim.autotune(params, edges=edges, trial_steps=8)
    assert snapshot(sim) == before
    trials = res["trials"]
    # 27 force-kernel candidates with the tile sort, then the best two with the counting sort
    assert len(trials) == 29 and not res["cached"]
    assert [t["sort_variant"] for t in trials] == [1] * 27 + [0] * 2
    timed = [t for t in trials if not t["pruned"]]
    best = min(timed, key=lambda t: t["us_per_step"])
    assert res["us_per_step"] == pytest.approx(best["us_per_step"], rel=1e-3)
    assert all(res["tuning"][k] == best[k] for k in ("cell_divisor", "force_variant",
                                                     "sort_variant"))
    assert sim.getTuning() == res["tuning"]
    print("chosen", res["tuning"], f"{res['us_per_step']:.2f} us/step")

    # a fresh simulation given the same state and the chosen tuning runs identically
    ref_state = make_sim(st, phys)
    ref_state.setParticles(sim.getParticleData())
    ref_state.setNeighborCounts(sim.getNeighborCounts())
    ref_state.setTuning(**res["tuning"])
    sim.simulate(params, 5, edges=edges)
    ref_state.simulate(params, 5, edges=edges)
    assert snapshot(sim) == snapshot(ref_state)
    if edges:
        assert np.array_equal(sim.getEdges(), ref_state.getEdges())


def test_tuning_key_leaves_state_and_is_stable():
    st, phys, params = state_of(small_cfg(n=5000, side=0.5))
    sim = make_sim(st, phys)
    sim.setTuning(cell_divisor=2)
    sim.simulate(params, 2)
    before = snapshot(sim)
    k1 = sim.tuningKey(params)
    k2 = sim.tuningKey(params, edges=(0.02, 20))
    assert snapshot(sim) == before
    assert sim.tuningKey(params) == k1 and k1.startswith("gfx")
    assert k1.endswith("/e=0") and k2.endswith("/e=32")
    assert "/T=6/" in k1 and "/n=12/" in k1 and "/sims=0/" in k1


def test_cache_hit_miss_and_force(tmp_path):
    path = tmp_path / "tune.json"
    st, phys, params = state_of(small_cfg(n=4000, side=0.5))
    a = make_sim(st, phys)
    r1 = a.autotune(params, trial_steps=4, cache=path)
    assert not r1["cached"]
    doc = json.load(open(path))
    assert list(doc["entries"]) == [r1["key"]]
    b = make_sim(st, phys)
    b.setTuning(steps_per_graph=8)
    r2 = b.autotune(params, trial_steps=4, cache=path)
    assert r2["cached"] and r2["trials"] == []
    expect = dict(r1["tuning"], steps_per_graph=8)  # the caller's graph length is kept
    assert r2["tuning"] == expect and b.getTuning() == expect
    r3 = b.autotune(params, trial_steps=4, cache=path, force=True)
    assert not r3["cached"] and len(r3["trials"]) == 29


def test_ensemble_autotune_leaves_state():
    cfg = dict(configs.load("c4_ensemble_256"), n_min=300, n_max=3000)
    sims = configs.make_ensemble(cfg, seed=5, sims=5)
    ens = Ensemble(sims)
    ens.simulate(2)
    before = [ens.getParticleData(s).tobytes() for s in range(5)]
    res = ens.autotune(trial_steps=4)
    assert [ens.getParticleData(s).tobytes() for s in range(5)] == before
    assert ens.getTuning() == res["tuning"] and "/sims=2/" in res["key"]
    ens.simulate(3)
    # every member still equals its plain run: default tuning for 2 steps, then the chosen one
    for s, (ph, st) in enumerate(sims):
        plain = make_sim(st, ph)
        prm = SimulationParams.from_physics(ph)
        plain.simulate(prm, 2)
        plain.setTuning(**res["tuning"])
        plain.simulate(prm, 3)
        assert ens.getParticleData(s).tobytes() == plain.getParticleData().tobytes(), s
        plain.close()
    ens.close()
