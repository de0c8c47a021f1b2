"""Shared checks for the GPU parity tests (engine vs oracle)."""

import numpy as np

from cellflow_amd import ParticleSimulation, SimulationParams, configs, particles_from_state

# Tolerances of one FP32 engine step against the FP64 oracle from the same FP32 input.
# Velocity: each pair term carries a few FP32 ulps (rsqrt, exp2, products), so the error of
# the force sum is bounded by a small multiple of 2^-24 * sum_j |net_ij| ("fabs"); 2e-5 of
# the magnitude scale leaves a wide margin.  Positions add the FP32 rounding of the wrap.
VEL_RTOL = 2e-5


def make_sim(st, phys):
    sim = ParticleSimulation(len(st["type"]), phys["num_types"])
    sim.setForceTable(st["table"])
    sim.setRadioByType(st["rbt"])
    sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
    if np.any(st["prev"]):
        sim.setNeighborCounts(st["prev"])
    return sim


def vel_scale(phys, vel_in, fabs):
    fmax = phys["force_multiplier"] * max(1.0, abs(phys["balance"]))
    return (np.abs(np.asarray(vel_in, np.float64)).max(axis=1) * abs(phys["friction"])
            + phys["delta_t"] * fmax * fabs + 1e-12)


def periodic_diff(a, b, L):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    L = np.asarray(L, np.float64)
    return d - L * np.round(d / L)


def check_step(got, ref, phys, vel_in):
    """got: engine records after one step; ref: oracle.step output on the same input."""
    L = np.asarray(phys["canvas"], np.float64)
    scale = vel_scale(phys, vel_in, ref["fabs"])
    dv = np.abs(got["vel"].astype(np.float64) - ref["vel"]).max(axis=1)
    rv = dv / scale
    ulp = np.spacing(np.float32(2 * L.max())).astype(np.float64)
    dx = np.abs(periodic_diff(got["pos"], ref["pos"], L)).max(axis=1)
    xtol = 4 * ulp + phys["delta_t"] * VEL_RTOL * scale
    da = np.abs(got["acc"].astype(np.float64) - ref["acc"]).max(axis=1)
    atol = VEL_RTOL * phys["force_multiplier"] * max(1.0, abs(phys["balance"])) * ref["fabs"] + 1e-9
    assert np.all(dv <= VEL_RTOL * scale), f"velocity error {rv.max():.3g} of scale at {rv.argmax()}"
    assert np.all(dx <= xtol), f"position error {np.max(dx / xtol):.3g} of tolerance"
    assert np.all(da <= atol), f"acceleration error {np.max(da / atol):.3g} of tolerance"
    return float(rv.max())


def small_cfg(n=3000, side=0.5, T=6, radius=0.05, ratio=0.3, seed=0, **over):
    cfg = dict(configs.load("c1_plain_10k"))
    cfg.update(n=n, canvas=[side, side, side], num_types=T, radius=radius, ratio=ratio, seed=seed)
    cfg.update(over)
    return cfg


def state_of(cfg, seed=None):
    return configs.make_state(cfg, seed), configs.physics(cfg), \
        SimulationParams.from_physics(configs.physics(cfg))


def check_step_edges(orc, sim, twin, params, edges, mode=0, steps=2):
    """The graph a step builds inside its force pass (the fused path, of the state before
    the step) against the oracle.  `sim` and `twin` hold the same state and tuning: `sim`
    runs `steps` steps with edges, so its last step builds the graph of the state after
    steps - 1; `twin` runs steps - 1 steps to that state (the engine is deterministic).
    Returns the number of edges."""
    sim.simulate(params, steps, edges=edges)
    twin.simulate(params, steps - 1, edges=edges)
    mid = twin.getParticleData()
    ref = orc.edges(mid["pos"], mid["ptype"], edges[0], edges[1], mode=mode)
    assert np.array_equal(twin.getEdges(), ref), "standalone graph"
    assert np.array_equal(sim.getStepEdges(), ref), "graph built inside the step"
    end = sim.getParticleData()
    assert np.array_equal(sim.getEdges(),
                          orc.edges(end["pos"], end["ptype"], edges[0], edges[1], mode=mode))
    return len(ref)
