"""Benchmark: particle-steps/s of the plain step at 100k particles with the proximity graph on.

Workload (cellflow_amd/configs/c2_100k_edges.json): 100,000 particles, 8 types, periodic
cube of side 2.0, force matrix N(0.1, 0.6), radius 0.04, ratio 0.3, density factor on,
edges within 0.02 capped at 16 per particle.  Inputs are synthetic (seeded SplitMix64)
and resident on the GPU before the timed region.  One timed step = binning + force +
integrate + stable re-sort + the proximity graph of the post-step state.

Prints one JSON line.  `roofline` is for the dominant kernel (k_force), timed live with
HIP events on the engine's stream; `cpu_baseline` times the oracle (FP64 scalar C++
restatement, OpenMP) on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
# compulsory state traffic per particle-step: read pos 12 + vel 12 + type 4 + previous
# count 4, write pos 12 + vel 12 + count 4
STEP_BYTES = 60.0
EDGE_BYTES = 20.0         # + per particle for the edge pass
EDGE_PAIR_BYTES = 8.0     # + per emitted edge (an int32 pair)
METRIC = "particle-steps/s @100k; batched-ensemble & adjoint speedup vs plain step"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def force_kernel_name(t, edges=True):
    """The force kernel a tuning launches, as rocprofv3 names it."""
    e = "true" if edges else "false"
    r, g = t["cell_divisor"], t["lanes_per_particle"]
    tl = "true" if t.get("sort_variant", 1) == 1 else "false"
    return {0: f"k_force_grp<{e}, {r}, {g}, {tl}>", 1: f"k_force<{e}>",
            2: f"k_force_lds<{e}, {r}>", 3: f"k_force_list<{e}, {r}, {g}, {tl}>"}[t["force_variant"]]


def read_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if that
    summary is of the same kernel."""
    vals = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" in line and line.split(":", 1)[0] in ("kernel", "traffic_bytes_per_launch"):
                    k, v = line.split(":", 1)
                    vals[k] = v.strip()
    except OSError:
        return None
    if vals.get("kernel") != kernel or "traffic_bytes_per_launch" not in vals:
        return None
    return float(vals["traffic_bytes_per_launch"])


def cpu_baseline(cfg, budget_s=10.0):
    import numpy as np

    import oracle
    from cellflow_amd import configs
    st = configs.make_state(cfg)
    phys = configs.physics(cfg)
    e = cfg["edges"]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pos = st["pos"].astype(np.float64)
    vel = st["vel"].astype(np.float64)
    prev = st["prev"].copy()
    steps = 0
    t0 = time.perf_counter()
    while True:
        out = oracle.step(pos, vel, st["type"], prev, phys, st["table"], st["rbt"], mode=1,
                          nthreads=threads)
        pos, vel, prev = out["pos"], out["vel"], out["count"]
        oracle.edges(pos.astype(np.float32), st["type"], e["distance"], e["max_connections"],
                     mode=1, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 5000:
            break
    return {"value": cfg["n"] * steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the 100k workload (grid enumeration, FP64, edges on), "
                      f"{el:.1f} s"}


def ensemble_leg(steps, warm=3):
    """Batched ensemble (256 simulations, uneven sizes) vs the same simulations run one after
    another as plain runs.  Both sides are warmed first (grids built, graphs captured)."""
    from cellflow_amd import Ensemble, ParticleSimulation, SimulationParams, particles_from_state
    from cellflow_amd import configs
    cfg = configs.load("c4_ensemble_256")
    sims = configs.make_ensemble(cfg)
    total = sum(len(st["type"]) for _, st in sims)
    log(f"[bench] ensemble: {len(sims)} simulations, {total} particles, {steps} steps")
    cache = os.path.join(tempfile.mkdtemp(prefix="cf_tune_"), "tuning.json")
    ens = Ensemble(sims)
    tuned = ens.autotune(cache=cache)
    log(f"[bench] ensemble tuning {tuned['tuning']} ({tuned['us_per_step']:.1f} us/step)")
    ens.simulate(warm)
    ens.getNeighborCounts(0)
    t0 = time.perf_counter()
    ens.simulate(steps)
    ens.getNeighborCounts(0)
    t_ens = time.perf_counter() - t0
    ens.close()
    plain = []
    for ph, st in sims:
        p = ParticleSimulation(len(st["type"]), ph["num_types"])
        p.setForceTable(st["table"])
        p.setRadioByType(st["rbt"])
        p.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
        prm = SimulationParams.from_physics(ph)
        p.autotune(prm, cache=cache)  # plain runs are tuned per workload bucket too
        p.simulate(prm, warm)
        plain.append((p, prm))
    plain[-1][0].getNeighborCounts()
    log("[bench] ensemble: sequential plain runs")
    t0 = time.perf_counter()
    for p, prm in plain:
        p.simulate(prm, steps)
    t_seq = time.perf_counter() - t0
    for p, _ in plain:
        p.close()
    return {"sims": len(sims), "particles": total, "steps": steps, "ensemble_s": t_ens,
            "sequential_s": t_seq, "speedup": t_seq / t_ens,
            "ensemble_particle_steps_per_s": total * steps / t_ens,
            "sequential_particle_steps_per_s": total * steps / t_seq,
            "workload": cfg["name"], "tuning": tuned["tuning"]}


def clustered_leg():
    """The 1,000,000-particle clustered workload (200 Gaussian blobs, radius 0.04, ratio 0.5,
    balance 1.2, edges 0.02/16) for its full 500 steps after autotuning: the regime where
    cells hold hundreds of particles and the pair loop dominates."""
    from cellflow_amd import ParticleSimulation, SimulationParams, configs, particles_from_state
    cfg = configs.load("c3_1m_clustered")
    st, ph = configs.make_state(cfg), configs.physics(cfg)
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    prm = SimulationParams.from_physics(ph)
    sim = ParticleSimulation(cfg["n"], cfg["num_types"])
    sim.setForceTable(st["table"])
    sim.setRadioByType(st["rbt"])
    sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
    tuned = sim.autotune(prm, edges=edges, trial_steps=4)
    sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
    log(f"[bench] clustered 1M: tuning {tuned['tuning']}; {cfg['steps']} steps")
    t0 = time.perf_counter()
    sim.simulate(prm, cfg["steps"], edges=edges)
    dt = time.perf_counter() - t0
    counts = sim.getNeighborCounts()
    out = {"workload": cfg["name"], "particles": cfg["n"], "steps": cfg["steps"],
           "ms_per_step": dt / cfg["steps"] * 1e3,
           "particle_steps_per_s": cfg["n"] * cfg["steps"] / dt, "tuning": tuned["tuning"],
           "mean_neighbours_final": float(counts.mean()), "edges_final": len(sim.getEdges())}
    sim.close()
    return out


def frame_leg(sim, params, edges, T, frames=200, warm=5):
    """CellFlowWidget's frame (CellFlowWidget.cpp:410-427, :551-575): one step with the
    proximity graph, then the particle records and the graph's line vertices copied to host
    memory.  The host-inclusive rate of the drop-in as the widget drives it; the headline
    keeps the state resident."""
    import numpy as np
    colors = np.random.default_rng(0).random((T, 3)).astype(np.float32)
    # host buffers reused from frame to frame, as the widget reuses its vectors
    pbuf = sim.getParticleData()
    vbuf = np.zeros(2 * sim.getLineVertices(colors).size + 1024, np.float32)
    for _ in range(warm):
        sim.simulate(params, 1, edges=edges)
        sim.getParticleData(out=pbuf)
        sim.getLineVertices(colors, out=vbuf)
    t0 = time.perf_counter()
    nbytes = 0
    for _ in range(frames):
        sim.simulate(params, 1, edges=edges)
        parts = sim.getParticleData(out=pbuf)
        verts = sim.getLineVertices(colors, out=vbuf)
        nbytes += parts.nbytes + verts.nbytes
    dt = time.perf_counter() - t0
    n = sim.getParticleCount()
    return {"frames": frames, "ms_per_frame": dt / frames * 1e3, "frames_per_s": frames / dt,
            "particle_steps_per_s": n * frames / dt, "host_bytes_per_frame": nbytes // frames,
            "note": "step + graph, records and line vertices to host each frame (PCIe-inclusive)"}


def mesh_init_leg(sim, T, reps=20):
    """Per-call times of the triangle mesh (generateTriangleMesh, ParticleSimulation.cu:
    689-754) on the current 100k state and of its vertex expansion, and of the seeded
    initialisation of 100k particles (the replacement for ParticleSimulation.cu:21-66)."""
    import numpy as np
    from cellflow_amd import ParticleSimulation
    colors = np.random.default_rng(0).random((T, 3)).astype(np.float32)
    tris = sim.generateTriangleMesh(0.02, 16)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.generateTriangleMesh(0.02, 16)
    mesh_ms = (time.perf_counter() - t0) / reps * 1e3
    sim.getTriangleVertices(colors)
    t0 = time.perf_counter()
    for _ in range(reps):
        v = sim.getTriangleVertices(colors)
    vert_ms = (time.perf_counter() - t0) / reps * 1e3
    n = sim.getParticleCount()
    fresh = ParticleSimulation(n, T)
    fresh.initializeParticles(1, (2.0, 2.0, 2.0), 2.0)
    t0 = time.perf_counter()
    for k in range(reps):
        fresh.initializeParticles(k, (2.0, 2.0, 2.0), 2.0)
    init_ms = (time.perf_counter() - t0) / reps * 1e3
    fresh.close()
    return {"triangle_mesh": {"distance": 0.02, "max_connections": 16, "triangles": len(tris),
                              "ms_per_build": mesh_ms, "vertices_ms": vert_ms,
                              "vertex_bytes": int(v.nbytes)},
            "seeded_init": {"particles": n, "ms_per_call": init_ms},
            "note": "per call, host-synchronous: triangles stay on the device until read; "
                    "the vertex expansion includes the copy to host"}


def adjoint_leg():
    """Differentiable rollout (20k particles, 200 steps, checkpoints every 20): time of the
    gradient rollout against the forward rollout; central differences over the 37
    parameters would need 74 forward rollouts."""
    from cellflow_amd import ParticleSimulation, SimulationParams, configs, particles_from_state
    cfg = configs.load("c5_adjoint_20k")

    def make(seed):
        st = configs.make_state(cfg, seed)
        ph = configs.physics(cfg)
        sim = ParticleSimulation(cfg["n"], cfg["num_types"])
        sim.setForceTable(st["table"])
        sim.setRadioByType(st["rbt"])
        sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
        return sim, st, SimulationParams.from_physics(ph)

    steps, ck, bins, rmax = cfg["steps"], cfg["checkpoint_every"], cfg["bins"], cfg["r_max"]
    tsim, _, tprm = make(cfg["target_seed"])
    tsim.simulate(tprm, steps)
    target = tsim.pairHistogram(bins, rmax)
    tsim.close()
    sim, st, prm = make(cfg["seed"])
    parts = particles_from_state(st["pos"], st["vel"], st["type"])
    sim.simulate(prm, 2)  # warm: grid, graphs
    sim.rolloutGrad(prm, 20, 10, bins, rmax, target)
    t_fwd, t_grad = [], []
    for _ in range(3):
        sim.setParticles(parts)
        sim.getNeighborCounts()
        t0 = time.perf_counter()
        sim.simulate(prm, steps)
        sim.pairHistogram(bins, rmax)
        t_fwd.append(time.perf_counter() - t0)
        sim.setParticles(parts)
        sim.getNeighborCounts()
        t0 = time.perf_counter()
        loss, _, _, _ = sim.rolloutGrad(prm, steps, ck, bins, rmax, target)
        t_grad.append(time.perf_counter() - t0)
    sim.close()
    tf, tg = min(t_fwd), min(t_grad)
    nparam = cfg["num_types"] ** 2 + 1
    return {"workload": cfg["name"], "particles": cfg["n"], "steps": steps,
            "checkpoint_every": ck, "forward_s": tf, "gradient_s": tg, "cost_ratio": tg / tf,
            "fd_rollouts": 2 * nparam, "speedup_vs_fd": 2 * nparam * tf / tg, "loss": loss}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout leg")
    ap.add_argument("--no-clustered", action="store_true", help="skip the 1M clustered leg")
    ap.add_argument("--no-frame", action="store_true", help="skip the host-inclusive frame leg")
    ap.add_argument("--no-mesh", action="store_true", help="skip the triangle-mesh/init leg")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (rocprofv3 counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.calibrate_copy} bytes 5 times")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this engine "
              "runs each simulation or ensemble on one GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cellflow_amd import ParticleSimulation, SimulationParams, configs, particles_from_state

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    cfg = configs.load("c2_100k_edges")
    st = configs.make_state(cfg)
    phys = configs.physics(cfg)
    params = SimulationParams.from_physics(phys)
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    n = cfg["n"]

    sim = ParticleSimulation(n, cfg["num_types"])
    sim.setForceTable(st["table"])
    sim.setRadioByType(st["rbt"])
    sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
    tuned = sim.autotune(params, edges=edges, trial_steps=64)
    log(f"[bench] tuning {tuned['tuning']} ({tuned['us_per_step']:.1f} us/step in trials)")
    log(f"[bench] warmup {args.warmup} steps")
    sim.simulate(params, args.warmup, edges=edges)
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    sim.simulate(params, args.steps, edges=edges)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nedges = len(sim.getEdges())
    rate = n * args.steps / dt

    log(f"[bench] {rate:.4g} particle-steps/s; per-kernel timing over {args.profile_steps} steps")
    kt = sim.profileSteps(params, args.profile_steps, edges=edges)
    force_ms = kt["force"]
    alg_bytes = n * (STEP_BYTES + EDGE_BYTES) + EDGE_PAIR_BYTES * nedges
    achieved = alg_bytes / (force_ms * 1e-3) / 1e9
    kname = force_kernel_name(tuned["tuning"])
    roof = {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": read_traffic(os.path.join(ROOT, "profiles", "k_force_c2_traffic.txt"),
                                    kname),
            "algorithmic_bytes_per_launch": alg_bytes, "avg_launch_ms": force_ms,
            "kernel_ms_per_launch": {k: round(v, 5) for k, v in kt.items()},
            "note": "the force kernel runs every step; the standalone edge pass (edges, "
                    "edge_scan, edge_write) runs once per call, after the last step"}

    out = {"metric": METRIC, "value": rate, "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": cfg["name"], "particles": n, "types": cfg["num_types"],
                      "box": cfg["canvas"][0], "radius": cfg["radius"], "ratio": cfg["ratio"],
                      "edges": e, "edges_last_step": nedges, "grid": sim.getGrid()[0]},
           "roofline": roof,
           "autotune": {"tuning": tuned["tuning"], "key": tuned["key"],
                        "us_per_step": tuned["us_per_step"],
                        "trials": [{k: t[k] for k in ("cell_divisor", "force_variant",
                                                      "lanes_per_particle", "block_size",
                                                      "us_per_step")} for t in tuned["trials"]]}}
    if not args.no_frame:
        log("[bench] frame loop (host-inclusive)")
        out["frame_loop"] = frame_leg(sim, params, edges, cfg["num_types"])
    if not args.no_mesh:
        log("[bench] triangle mesh and seeded init")
        out["mesh_and_init"] = mesh_init_leg(sim, cfg["num_types"])
    sim.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ensemble_steps)
    if not args.no_adjoint:
        log("[bench] differentiable rollout")
        out["adjoint"] = adjoint_leg()
    if not args.no_clustered:
        out["clustered"] = clustered_leg()
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
