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
            2: f"k_force_lds<{e}, {r}>", 3: f"k_force_list<{e}, {r}, {g}, {tl}, false>",
            4: f"k_force_list<{e}, {r}, {g}, {tl}, true>"}[t["force_variant"]]


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


def _oracle_rate(oracle, cfg, steps, mode, threads, edges=None, keep=0):
    """Particle-steps/s of the oracle over `steps` steps of a workload from its seeded start
    (FP64, mode 0 all-pairs as the reference enumerates, 1 grid), with the graph of every
    post-step state when `edges` is given; also the first `keep` states."""
    import numpy as np

    from cellflow_amd import configs
    st, phys = configs.make_state(cfg), configs.physics(cfg)
    pos, vel, prev = st["pos"].astype(np.float64), st["vel"].astype(np.float64), st["prev"].copy()
    kept = []
    t0 = time.perf_counter()
    for k in range(steps):
        out = oracle.step(pos, vel, st["type"], prev, phys, st["table"], st["rbt"], mode=mode,
                          nthreads=threads)
        pos, vel, prev = out["pos"], out["vel"], out["count"]
        if edges:
            oracle.edges(pos.astype(np.float32), st["type"], edges[0], edges[1], mode=mode,
                         nthreads=threads)
        if k < keep:
            kept.append((pos, vel, prev))
    el = time.perf_counter() - t0
    return cfg["n"] * steps / el, el, kept


def cpu_baseline(cfg, params, budget_s=10.0):
    """The CPU-baseline leg.  The oracle (FP64 scalar C++ restatement of the reference step)
    on the host cores: the 100k workload with edges on all threads for about `budget_s`
    seconds (the reported value), plus single-thread and 10k-workload samples in both
    enumerations.  The oracle's first 10 states of the 100k sample then check a 10-step
    engine rollout from the same start (parity_sample).  Returns (baseline, parity)."""
    import oracle
    from cellflow_amd import configs
    e = cfg["edges"]
    edges = (e["distance"], e["max_connections"])
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # the main sample: as many steps as fit the budget (calibrated on two steps)
    r2, el2, _ = _oracle_rate(oracle, cfg, 2, 1, threads, edges)
    steps = max(10, min(5000, int(budget_s / (el2 / 2))))
    rate, el, kept = _oracle_rate(oracle, cfg, steps, 1, threads, edges, keep=10)
    one, el1, _ = _oracle_rate(oracle, cfg, 5, 1, 1, edges)
    c1 = configs.load("c1_plain_10k")
    samples = {
        "c2_grid_all_threads": {"value": rate, "steps": steps, "s": el, "threads": threads},
        "c2_grid_1_thread": {"value": one, "steps": 5, "s": el1, "threads": 1},
    }
    for mode, tag, n_all, n_one in ((1, "grid", c1["steps"], c1["steps"]), (0, "all_pairs", 10, 2)):
        ra, ea, _ = _oracle_rate(oracle, c1, n_all, mode, threads)
        r1, e1, _ = _oracle_rate(oracle, c1, n_one, mode, 1)
        samples[f"c1_{tag}_all_threads"] = {"value": ra, "steps": n_all, "s": ea, "threads": threads}
        samples[f"c1_{tag}_1_thread"] = {"value": r1, "steps": n_one, "s": e1, "threads": 1}
    base = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the 100k workload (grid enumeration, FP64, edges on), "
                      f"{el:.1f} s on {threads} threads",
            "single_thread": one, "samples": samples,
            "note": "10k samples: the plain 10k workload (no edges); all_pairs enumerates every "
                    "pair as the reference kernel does (ParticleSimulation.cu:86-89), grid "
                    "uses the cell stencil (same neighbour set)"}
    return base, parity_sample(cfg, params, kept)


def parity_sample(cfg, params, kept):
    """Part of the CPU-baseline leg (the oracle is the checker): the engine's first 10
    steps of the 100k workload (edges on) against the oracle's FP64 states from the CPU
    baseline's sample, from the same FP32 start.  Errors are over
    the particles outside the light cone of every decision flip (count mismatch): a flip
    marks both particles and each step marks every particle in range of a marked one.
    pos: max |x - x64| / L (periodic); vel: max |v - v64| / max |v64|."""
    import numpy as np

    import oracle
    from cellflow_amd import ParticleSimulation, configs, particles_from_state
    st, phys = configs.make_state(cfg), configs.physics(cfg)
    e = cfg["edges"]
    L = np.asarray(phys["canvas"], np.float64)
    sim = ParticleSimulation(cfg["n"], cfg["num_types"])
    sim.setForceTable(st["table"])
    sim.setRadioByType(st["rbt"])
    sim.setParticles(particles_from_state(st["pos"], st["vel"], st["type"]))
    marked = np.zeros(cfg["n"], bool)
    prev_ref = st["pos"]
    rows = []
    for k, (pos64, vel64, cnt64) in enumerate(kept):
        cur = sim.getParticleData()
        for p in (prev_ref, cur["pos"]):
            offs, nbr = oracle.pairs(np.asarray(p, np.float32), st["type"], phys, st["rbt"])
            owner = np.repeat(np.arange(cfg["n"]), np.diff(offs))
            hit = np.zeros(cfg["n"], bool)
            np.logical_or.at(hit, owner, marked[nbr])
            marked |= hit
        sim.simulate(params, 1, edges=(e["distance"], e["max_connections"]))
        got = sim.getParticleData()
        flips = sim.getNeighborCounts() != cnt64
        marked |= flips
        clean = ~marked
        d = got["pos"].astype(np.float64) - pos64
        d -= L * np.round(d / L)
        ex = float(np.abs(d[clean]).max() / L.max())
        ev = float(np.abs(got["vel"].astype(np.float64) - vel64)[clean].max() / np.abs(vel64).max())
        rows.append({"step": k + 1, "flips": int(flips.sum()),
                     "flip_free_fraction": float(clean.mean()), "pos_rel": ex, "vel_rel": ev})
        prev_ref = pos64.astype(np.float32)
    sim.close()
    last = rows[-1]
    return {"workload": cfg["name"], "steps": len(rows), "reference": "oracle FP64, same FP32 start",
            "max_pos_rel": max(r["pos_rel"] for r in rows),
            "max_vel_rel": max(r["vel_rel"] for r in rows),
            "first_step": rows[0], "last_step": last, "per_step": rows,
            "note": "free-running; errors over the particles outside the light cone of "
                    "every decision flip (the dynamics are chaotic: an FP64 run started one "
                    "FP32 rounding away diverges by a similar amount)"}


def ensemble_leg(steps, warm=3, tuning="", sequential=True):
    """Batched ensemble (256 simulations, uneven sizes) vs the same simulations run one after
    another as plain runs.  Both sides are warmed first (grids built, graphs captured).
    `tuning` (force_variant,lanes,cell_divisor,sort_variant) replaces the ensemble's autotuning
    and `sequential` False skips the plain runs (profiling runs of the ensemble alone)."""
    from cellflow_amd import Ensemble, ParticleSimulation, SimulationParams, particles_from_state
    from cellflow_amd import configs
    cfg = configs.load("c4_ensemble_256")
    sims = configs.make_ensemble(cfg)
    total = sum(len(st["type"]) for _, st in sims)
    log(f"[bench] ensemble: {len(sims)} simulations, {total} particles, {steps} steps")
    cache = os.path.join(tempfile.mkdtemp(prefix="cf_tune_"), "tuning.json")
    ens = Ensemble(sims)
    if tuning:
        v, g, d, so = (int(x) for x in tuning.split(","))
        ens.setTuning(cell_divisor=d, force_variant=v, lanes_per_particle=g, sort_variant=so)
        tuned = {"tuning": ens.getTuning(), "us_per_step": float("nan")}
    else:
        tuned = ens.autotune(cache=cache)
    log(f"[bench] ensemble tuning {tuned['tuning']} ({tuned['us_per_step']:.1f} us/step)")
    ens.simulate(warm)
    ens.getNeighborCounts(0)
    t0 = time.perf_counter()
    ens.simulate(steps)
    ens.getNeighborCounts(0)
    t_ens = time.perf_counter() - t0
    ens.close()
    if not sequential:
        return {"sims": len(sims), "particles": total, "steps": steps, "ensemble_s": t_ens,
                "ensemble_particle_steps_per_s": total * steps / t_ens,
                "workload": cfg["name"], "tuning": tuned["tuning"]}
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


def frame_gpu_leg(sim, params, edges, frames=50):
    """GPU time of the widget's per-frame work (CellFlowWidget.cpp:424, :569): one step,
    then the proximity graph of the post-step state by the standalone pass, per kernel
    (HIP events on the engine's stream), averaged over `frames` frames."""
    acc = {}
    for _ in range(frames):
        for k, v in sim.profileSteps(params, 1, edges=edges).items():
            acc[k] = acc.get(k, 0.0) + v
    per = {k: v / frames for k, v in acc.items()}
    return {"frames": frames, "kernel_ms": {k: round(v, 5) for k, v in per.items()},
            "gpu_ms_per_frame": sum(per.values()),
            "graph_ms": sum(v for k, v in per.items() if k.startswith("edge")),
            "note": "simulate(1 step, edges): the step's kernels, then the standalone graph "
                    "pass of the post-step state, as the widget drives it once per frame"}


def read_counters(path, kernel):
    """Per-launch counter values of `kernel` from a committed rocprofv3 PMC summary
    ('counter: value' lines after a 'kernel: <name>' line), if the summary is of that
    kernel."""
    vals, cur = {}, None
    try:
        with open(path) as f:
            for line in f:
                if ":" not in line or line.startswith((" ", "|", "#")):
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "kernel":
                    cur = v
                elif cur == kernel:
                    try:
                        vals[k] = float(v)
                    except ValueError:
                        pass
    except OSError:
        return None
    return vals or None


# gfx950: 256 CUs; a CU issues one wave64 VALU instruction per cycle (4 SIMD16s, 4 cycles
# each); 2.4 GHz peak engine clock; FP32 vector peak 157.3 TFLOP/s
VALU_ISSUE_PEAK = 256 * 2.4e9
FP32_PEAK = 157.3e12


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
    ap.add_argument("--ensemble-tuning", default="",
                    help="force_variant,lanes,cell_divisor,sort_variant for the ensemble instead "
                         "of autotuning it (profiling runs)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="skip the ensemble's sequential plain runs (profiling runs)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout leg")
    ap.add_argument("--no-clustered", action="store_true", help="skip the 1M clustered leg")
    ap.add_argument("--no-frame", action="store_true", help="skip the host-inclusive frame leg")
    ap.add_argument("--no-mesh", action="store_true", help="skip the triangle-mesh/init leg")
    ap.add_argument("--tuning", default="",
                    help="force_variant,lanes,cell_divisor,sort_variant: use this tuning instead "
                         "of autotuning (for profiling runs whose counters should hold the "
                         "timed steps only)")
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
    if args.tuning:
        v, g, d, so = (int(x) for x in args.tuning.split(","))
        sim.setTuning(cell_divisor=d, force_variant=v, lanes_per_particle=g, sort_variant=so)
        tuned = {"tuning": sim.getTuning(), "key": sim.tuningKey(params, edges),
                 "us_per_step": float("nan"), "trials": []}
    else:
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
                    "edge_emit) runs once per call, after the last step"}
    ctr = read_counters(os.path.join(ROOT, "profiles", "k_force_c2_valu.txt"), kname)
    if ctr and "SQ_INSTS_VALU" in ctr:
        import numpy as np
        # useful FP32 work: 12 FLOP per stencil candidate (difference, distance, cutoff)
        # plus 25 per in-range pair (the force term), on the state after the timed run
        dims, _ = sim.getGrid()
        cells, _ = sim.getSortOrder()
        occ = np.bincount(cells, minlength=int(np.prod(dims))).reshape(dims[::-1])
        stencil = sum(np.roll(occ, (dz, dy, dx), (0, 1, 2))
                      for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1))
        cand = float((occ * stencil).sum())
        inrange = float(sim.getNeighborCounts().sum())
        flops = 12.0 * cand + 25.0 * inrange
        vi = ctr["SQ_INSTS_VALU"]
        roof["valu"] = {"instr_per_launch": vi, "salu_instr_per_launch": ctr.get("SQ_INSTS_SALU"),
                        "issue_us": vi / VALU_ISSUE_PEAK * 1e6,
                        "issue_frac": vi / VALU_ISSUE_PEAK / (force_ms * 1e-3),
                        "useful_flop_per_launch": flops,
                        "flop_frac": flops / (force_ms * 1e-3) / FP32_PEAK,
                        "source": "profiles/k_force_c2_valu.txt (rocprofv3 --pmc)",
                        "note": "issue_frac: VALU wave instructions at one per CU per cycle "
                                "(2.4 GHz) over the launch time; flop_frac: 12 FLOP per "
                                "27-cell stencil candidate + 25 per in-range pair over FP32 "
                                "peak"}

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
        log("[bench] widget frame: GPU time per kernel")
        out["frame_gpu"] = frame_gpu_leg(sim, params, edges)
        log("[bench] frame loop (host-inclusive)")
        out["frame_loop"] = frame_leg(sim, params, edges, cfg["num_types"])
    if not args.no_mesh:
        log("[bench] triangle mesh and seeded init")
        out["mesh_and_init"] = mesh_init_leg(sim, cfg["num_types"])
    sim.close()
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ensemble_steps, tuning=args.ensemble_tuning,
                                       sequential=not args.no_sequential)
    if not args.no_adjoint:
        log("[bench] differentiable rollout")
        out["adjoint"] = adjoint_leg()
    if not args.no_clustered:
        out["clustered"] = clustered_leg()
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) and a 10-step parity sample against it")
        out["cpu_baseline"], out["parity"] = cpu_baseline(cfg, params)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
