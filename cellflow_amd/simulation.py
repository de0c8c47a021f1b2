"""Python mirror of the reference's ``ParticleSimulation`` class over the C ABI.

Method names, argument meanings and defaults follow
cuda-native/include/ParticleSimulation.cuh:10-103 and SimulationParams.h:14-57, so code
written against the reference reads the same here.  Where the reference calls exit(1)
(CUDA_CHECK, ParticleSimulation.cu:10-18) this raises :class:`CellFlowError`.
Arrays are numpy (host memory); device memory stays inside the engine.
"""

from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import CellFlowError, check, lib


@dataclass
class SimulationParams:
    """The step fields of SimulationParams (SimulationParams.h:14-29) with their defaults."""
    radius: float = 42.07
    delta_t: float = 0.18
    friction: float = 0.51
    repulsion: float = 64.83
    attraction: float = 3.06
    k: float = 29.45
    balance: float = 0.79
    canvasWidth: float = 8000.0
    canvasHeight: float = 8000.0
    canvasDepth: float = 8000.0
    numParticleTypes: int = 6
    ratioWithLFO: float = 0.0
    forceMultiplier: float = 2.33
    maxExpectedNeighbors: int = 400

    @classmethod
    def from_physics(cls, phys: dict) -> "SimulationParams":
        """From a :func:`cellflow_amd.configs.physics` dict."""
        L = phys["canvas"]
        return cls(radius=phys["radius"], delta_t=phys["delta_t"], friction=phys["friction"],
                   repulsion=phys["repulsion"], attraction=phys["attraction"], k=phys["k"],
                   balance=phys["balance"], canvasWidth=L[0], canvasHeight=L[1],
                   canvasDepth=L[2], numParticleTypes=phys["num_types"],
                   ratioWithLFO=phys["ratio"], forceMultiplier=phys["force_multiplier"],
                   maxExpectedNeighbors=phys["max_expected_neighbors"])

    def to_c(self) -> _abi.cf_params:
        p = _abi.cf_params()
        p.radius, p.delta_t, p.friction = self.radius, self.delta_t, self.friction
        p.repulsion, p.attraction, p.k, p.balance = (self.repulsion, self.attraction, self.k,
                                                     self.balance)
        p.canvas_width, p.canvas_height, p.canvas_depth = (self.canvasWidth, self.canvasHeight,
                                                           self.canvasDepth)
        p.num_particle_types = self.numParticleTypes
        p.ratio_with_lfo = self.ratioWithLFO
        p.force_multiplier = self.forceMultiplier
        p.max_expected_neighbors = self.maxExpectedNeighbors
        return p


PARTICLE_DTYPE = np.dtype([("pos", np.float32, 3), ("vel", np.float32, 3),
                           ("acc", np.float32, 3), ("ptype", np.uint32), ("pad", np.float32)])
assert PARTICLE_DTYPE.itemsize == 44


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a, ct=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ct))


def _reuse(out, count, dtype):
    """A flat view of `out`'s first `count` elements, to be filled in place, or a new zeroed
    array when `out` is None.  `out` must be a writeable C-contiguous array of `dtype` (any
    shape) holding at least `count` elements: anything else raises ValueError instead of
    being silently replaced by a copy the caller never sees."""
    if out is None:
        return np.zeros(count, dtype)
    if not isinstance(out, np.ndarray) or out.dtype != dtype:
        raise ValueError(f"out must be a numpy array of dtype {dtype}")
    if not (out.flags.c_contiguous and out.flags.writeable):
        raise ValueError("out must be a writeable C-contiguous array")
    if out.size < count:
        raise ValueError(f"out holds {out.size} elements, {count} needed")
    return out.reshape(-1)[:count]


class ParticleSimulation:
    """Owner of one simulation's device state (ParticleSimulation.cuh:10-103)."""

    def __init__(self, particleCount: int, numParticleTypes: int = 6):
        h = C.c_void_p()
        check(lib().cf_create(int(particleCount), int(numParticleTypes), C.byref(h)))
        self._h = h
        self._edge_opts = None

    # -- lifetime ---------------------------------------------------------------------
    def close(self):
        if self._h:
            check(lib().cf_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().cf_destroy(self._h)
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- counts and types (ParticleSimulation.cuh:39-43) --------------------------------
    def getParticleCount(self) -> int:
        n = C.c_int()
        check(lib().cf_get_particle_count(self._h, C.byref(n)))
        return n.value

    def setParticleCount(self, count: int):
        check(lib().cf_set_particle_count(self._h, int(count)))

    def setNumParticleTypes(self, types: int):
        check(lib().cf_set_num_particle_types(self._h, int(types)))

    # -- state ----------------------------------------------------------------------------
    def setParticles(self, particles: np.ndarray):
        """Load index-ordered 44-byte records (PARTICLE_DTYPE); counts reset to zero."""
        a = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        check(lib().cf_set_particles(self._h, a.ctypes.data_as(C.POINTER(_abi.cf_particle)),
                                     len(a)))

    def getParticleData(self, out: np.ndarray = None) -> np.ndarray:
        """getParticleData (ParticleSimulation.cu:558-562): records in index order.  `out`, a
        contiguous array of PARTICLE_DTYPE with at least the particle count of records, is
        filled and returned (sliced to the count) instead of a new array, as the reference
        reuses the caller's vector."""
        n = self.getParticleCount()
        out = _reuse(out, n, PARTICLE_DTYPE)
        check(lib().cf_get_particles(self._h, out.ctypes.data_as(C.POINTER(_abi.cf_particle)), n))
        return out

    def setState(self, pos, type_, vel=None):
        pos = np.ascontiguousarray(pos, np.float32)
        t = np.ascontiguousarray(type_, np.uint32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        check(lib().cf_set_state_soa(self._h, _fp(pos), None if v is None else _fp(v),
                                     _ip(t, C.c_uint32), len(t)))

    def getState(self):
        n = self.getParticleCount()
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        t = np.zeros(n, np.uint32)
        check(lib().cf_get_state_soa(self._h, _fp(pos), _fp(vel), _ip(t, C.c_uint32), n))
        return pos, vel, t

    def setNeighborCounts(self, counts):
        c = np.ascontiguousarray(counts, np.int32)
        check(lib().cf_set_neighbor_counts(self._h, _ip(c), len(c)))

    def getNeighborCounts(self) -> np.ndarray:
        n = self.getParticleCount()
        c = np.zeros(n, np.int32)
        check(lib().cf_get_neighbor_counts(self._h, _ip(c), n))
        return c

    def initializeParticles(self, seed: int, canvas=(8000.0, 8000.0, 8000.0),
                            spawnRegionSize: float = 2000.0):
        """Seeded initializeParticles (ParticleSimulation.cu:484-492, spawn cube 2000)."""
        cv = np.asarray(canvas, np.float32)
        check(lib().cf_init_particles(self._h, int(seed), _fp(cv), float(spawnRegionSize)))

    def moveUniverse(self, dx: float, dy: float, dz: float = 0.0,
                     canvas=(8000.0, 8000.0, 8000.0)):
        """moveUniverse (ParticleSimulation.cu:585-592): x = fmodf(x + d + L, L) per axis.
        The reference always uses a default SimulationParams' canvas (8000^3) and dz = 0."""
        cv = np.asarray(canvas, np.float32)
        check(lib().cf_move_universe(self._h, float(dx), float(dy), float(dz), _fp(cv)))

    # -- preset files (CellFlowWidget.cpp:1070-1269) ----------------------------------------
    _PARAM_KEYS = {"radius": "radius", "delta_t": "delta_t", "friction": "friction",
                   "repulsion": "repulsion", "attraction": "attraction", "k": "k",
                   "balance": "balance", "forceMultiplier": "forceMultiplier",
                   "canvasWidth": "canvasWidth", "canvasHeight": "canvasHeight",
                   "canvasDepth": "canvasDepth"}

    def loadPreset(self, path, params: "SimulationParams | None" = None, seed: int = 1):
        """CellFlowWidget::loadPreset on this simulation, in the widget's order: particle
        count, type count, step parameters, radius modifiers (setRadioByTypeValue for the
        first min(len, T)), raw table entries (the first min(len, T*T)), then
        updateForceTable(forceRange, forceBias, forceOffset).  Fields the file lacks keep
        their values in `params` (default SimulationParams()).  When the count or type count
        changes the particles are re-initialised from `seed` in the preset's canvas (the
        reference re-initialises from a time seed).  Returns (params, preset dict)."""
        from dataclasses import replace

        from .presets import read_preset
        pr = read_preset(path)
        prm = replace(params) if params is not None else SimulationParams()
        reinit = False
        if "PARTICLE_COUNT" in pr and pr["PARTICLE_COUNT"] != self.getParticleCount():
            self.setParticleCount(pr["PARTICLE_COUNT"])
            reinit = True
        if "numParticleTypes" in pr:
            T = pr["numParticleTypes"]
            self.setNumParticleTypes(T)  # raw table and radii restart at zero (the reference
            reinit = True                 # draws random ones; the preset overwrites them)
        T = self.getNumParticleTypes()
        prm.numParticleTypes = T
        for key, attr in self._PARAM_KEYS.items():
            if key in pr:
                setattr(prm, attr, pr[key])
        if "ratio" in pr:
            prm.ratioWithLFO = pr["ratio"]  # LFO phase 0
        for i, v in enumerate(pr.get("radioByType", [])[:T]):
            self.setRadioByTypeValue(i, float(v))
        raw = self.getRawForceTableValues()
        vals = pr.get("rawForceTable", np.zeros(0, np.float32))[: T * T]
        raw[: len(vals)] = vals
        self.setRawForceTableValues(raw)
        self.updateForceTable(pr.get("forceRange", 0.28), pr.get("forceBias", -0.20),
                              pr.get("forceOffset", 1.0))
        if reinit:
            self.initializeParticles(seed, (prm.canvasWidth, prm.canvasHeight, prm.canvasDepth),
                                     pr.get("spawnRegionSize", 2000.0))
        return prm, pr

    def savePreset(self, path, params: "SimulationParams", preset: "dict | None" = None,
                   template=None):
        """CellFlowWidget::savePreset: the simulation's count, types, radius modifiers and
        raw table, the step parameters of `params`, and the UI fields of `preset` (force
        range/bias/offset, ratio, LFO, spawn region, colours).  With `template`, every other
        member of that file (render and camera settings) is kept."""
        from .presets import write_preset
        d = dict(preset or {})
        T = self.getNumParticleTypes()
        d["PARTICLE_COUNT"] = self.getParticleCount()
        d["numParticleTypes"] = T
        for key, attr in self._PARAM_KEYS.items():
            d[key] = getattr(params, attr)
        d.setdefault("ratio", params.ratioWithLFO)
        d["radioByType"] = self.getRadioByType()
        d["rawForceTable"] = self.getRawForceTableValues()
        write_preset(path, d, template)

    # -- force table and radii (ParticleSimulation.cuh:24-57) --------------------------------
    def setForceTable(self, table):
        t = np.ascontiguousarray(table, np.float32)
        T = int(round(np.sqrt(t.size)))
        check(lib().cf_set_force_table(self._h, _fp(t.ravel()), T))

    def getForceTable(self) -> np.ndarray:
        T = self.getNumParticleTypes()
        t = np.zeros(T * T, np.float32)
        check(lib().cf_get_force_table(self._h, _fp(t), T * T))
        return t.reshape(T, T)

    def setRawForceTableValues(self, raw):
        r = np.ascontiguousarray(raw, np.float32).ravel()
        check(lib().cf_set_raw_force_table(self._h, _fp(r), r.size))

    def getRawForceTableValues(self) -> np.ndarray:
        T = self.getNumParticleTypes()
        r = np.zeros(T * T, np.float32)
        check(lib().cf_get_raw_force_table(self._h, _fp(r), T * T))
        return r

    def updateForceTable(self, forceRange: float, forceBias: float, forceOffset: float):
        check(lib().cf_update_force_table(self._h, forceRange, forceBias, forceOffset))

    def setRadioByType(self, values):
        v = np.ascontiguousarray(values, np.float32)
        check(lib().cf_set_radio_by_type(self._h, _fp(v), len(v)))

    def setRadioByTypeValue(self, index: int, value: float):
        check(lib().cf_set_radio_by_type_value(self._h, int(index), float(value)))

    def getNumParticleTypes(self) -> int:
        t = C.c_int()
        check(lib().cf_get_num_particle_types(self._h, C.byref(t)))
        return t.value

    def getRadioByType(self) -> np.ndarray:
        """getRadioByType (ParticleSimulation.cu:607-613)."""
        T = self.getNumParticleTypes()
        v = np.zeros(T, np.float32)
        check(lib().cf_get_radio_by_type(self._h, _fp(v), T))
        return v

    def rotateRadioByType(self):
        check(lib().cf_rotate_radio_by_type(self._h))

    # -- the step and the proximity graph ---------------------------------------------------
    def simulate(self, params: SimulationParams, steps: int = 1, edges=None):
        """`steps` calls of simulate(params); edges=(distance, maxConnections) also builds the
        proximity graph of every post-step state (CellFlowWidget.cpp:551-575)."""
        p = params.to_c()
        eo = None
        if edges is not None:
            eo = _abi.cf_edge_opts(float(edges[0]), int(edges[1]))
        check(lib().cf_simulate(self._h, C.byref(p), int(steps),
                                None if eo is None else C.byref(eo)))

    def generateProximityGraph(self, proximityDistance: float, maxConnectionsPerParticle: int):
        """generateProximityGraph (ParticleSimulation.cu:624-687) on the current state;
        returns the (E, 2) int32 edge list in canonical (owner, rank) order."""
        check(lib().cf_build_edges(self._h, float(proximityDistance),
                                   int(maxConnectionsPerParticle)))
        return self.getEdges()

    def getEdges(self) -> np.ndarray:
        cnt = C.c_longlong()
        rc = lib().cf_get_edges(self._h, None, 0, C.byref(cnt))
        if rc not in (_abi.CF_OK, _abi.CF_ERR_CAPACITY):
            check(rc)
        out = np.zeros((max(cnt.value, 1), 2), np.int32)
        check(lib().cf_get_edges(self._h, _ip(out), cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def getStepEdges(self) -> np.ndarray:
        """The graph the last simulate call built inside its final step (that of the state
        before the final step), in the layout of :meth:`getEdges`; raises CF_ERR_STATE when
        that call built none (fewer than two steps, or no edges)."""
        cnt = C.c_longlong()
        rc = lib().cf_get_step_edges(self._h, None, 0, C.byref(cnt))
        if rc not in (_abi.CF_OK, _abi.CF_ERR_CAPACITY):
            check(rc)
        out = np.zeros((max(cnt.value, 1), 2), np.int32)
        check(lib().cf_get_step_edges(self._h, _ip(out), cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def generateTriangleMesh(self, proximityDistance: float, maxConnectionsPerParticle: int):
        """generateTriangleMesh (ParticleSimulation.cu:689-754) on the current state; returns
        the (T, 3) int32 triangles in canonical (owner, fan) order."""
        check(lib().cf_build_triangles(self._h, float(proximityDistance),
                                       int(maxConnectionsPerParticle)))
        return self.getTriangles()

    def getTriangles(self) -> np.ndarray:
        cnt = C.c_longlong()
        rc = lib().cf_get_triangles(self._h, None, 0, C.byref(cnt))
        if rc not in (_abi.CF_OK, _abi.CF_ERR_CAPACITY):
            check(rc)
        out = np.zeros((max(cnt.value, 1), 3), np.int32)
        check(lib().cf_get_triangles(self._h, _ip(out), cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def getTriangleVertices(self, colors, out: np.ndarray = None) -> np.ndarray:
        """The reference's triangle-vertex layout (ParticleSimulation.cu:362-418): (3T, 9)
        rows of position, flat normal and colour.  `out`, a writeable C-contiguous float32
        array, is filled in place when given (ValueError if it is too small)."""
        col = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
        vc = C.c_longlong()
        rc = lib().cf_get_triangle_vertices(self._h, col.ctypes.data_as(C.POINTER(_abi.cf_color)),
                                            len(col), None, 0, C.byref(vc))
        if rc not in (_abi.CF_OK, _abi.CF_ERR_CAPACITY):
            check(rc)
        out = _reuse(out, max(9 * vc.value, 1), np.float32)
        check(lib().cf_get_triangle_vertices(self._h,
                                             col.ctypes.data_as(C.POINTER(_abi.cf_color)),
                                             len(col), _fp(out), 9 * vc.value, C.byref(vc)))
        return out[: 9 * vc.value].reshape(-1, 9)

    def getLineVertices(self, colors, out: np.ndarray = None) -> np.ndarray:
        """The reference's GL line-vertex layout (ParticleSimulation.cu:253-275).  `out`, a
        writeable C-contiguous float32 array, is filled in place when given (ValueError if
        it is too small)."""
        col = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
        vc = C.c_longlong()
        rc = lib().cf_get_line_vertices(self._h, col.ctypes.data_as(C.POINTER(_abi.cf_color)),
                                        len(col), None, 0, C.byref(vc))
        if rc not in (_abi.CF_OK, _abi.CF_ERR_CAPACITY):
            check(rc)
        out = _reuse(out, max(6 * vc.value, 1), np.float32)
        check(lib().cf_get_line_vertices(self._h, col.ctypes.data_as(C.POINTER(_abi.cf_color)),
                                         len(col), _fp(out), 6 * vc.value, C.byref(vc)))
        return out[: 6 * vc.value].reshape(-1, 6)

    def _write_device(self, fn, colors, device_ptr: int, capacity_floats: int) -> int:
        col = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
        vc = C.c_longlong()
        check(fn(self._h, col.ctypes.data_as(C.POINTER(_abi.cf_color)), len(col),
                 C.c_void_p(int(device_ptr)), int(capacity_floats), C.byref(vc)))
        return vc.value

    def writeLineVerticesDevice(self, colors, device_ptr: int, capacity_floats: int) -> int:
        """generateProximityGraph's VBO write (ParticleSimulation.cu:253-275, :633-643): the
        line vertices of the last graph written by the GPU into caller-owned device memory
        at `device_ptr` (e.g. a mapped GL buffer, or a torch tensor's data_ptr()), holding
        `capacity_floats` floats.  Returns the vertex count (2 per edge)."""
        return self._write_device(lib().cf_write_line_vertices_device, colors, device_ptr,
                                  capacity_floats)

    def writeTriangleVerticesDevice(self, colors, device_ptr: int, capacity_floats: int) -> int:
        """generateTriangleMesh's VBO write (ParticleSimulation.cu:362-418): the triangle
        vertices of the last mesh into caller-owned device memory.  Returns the vertex
        count (3 per triangle)."""
        return self._write_device(lib().cf_write_triangle_vertices_device, colors, device_ptr,
                                  capacity_floats)

    # -- differentiable rollout ------------------------------------------------------------
    def pairHistogram(self, bins: int, r_max: float) -> np.ndarray:
        """Soft pair histogram of the current state (cubic B-spline bins, FP64)."""
        h = np.zeros(bins)
        check(lib().cf_pair_histogram(self._h, int(bins), float(r_max),
                                      h.ctypes.data_as(C.POINTER(C.c_double))))
        return h

    def rolloutGrad(self, params: SimulationParams, steps: int, checkpoint_every: int, bins: int,
                    r_max: float, target):
        """Roll `steps` steps from the current state and return (loss, histogram,
        dL/dtable (T x T), dL/dratio) of L = sum_b (h_b - target_b)^2 at the final state."""
        T = self.getNumParticleTypes()
        tgt = np.ascontiguousarray(target, np.float64)
        loss = C.c_double()
        h = np.zeros(bins)
        gt = np.zeros(T * T)
        gr = C.c_double()
        p = params.to_c()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        check(lib().cf_rollout_grad(self._h, C.byref(p), int(steps), int(checkpoint_every),
                                    int(bins), float(r_max), dp(tgt), C.byref(loss), dp(h),
                                    dp(gt), C.byref(gr)))
        return loss.value, h, gt.reshape(T, T), gr.value

    # -- introspection --------------------------------------------------------------------
    def getGrid(self):
        d = (C.c_int * 3)()
        r = C.c_int()
        check(lib().cf_get_grid(self._h, d, C.byref(r)))
        return (d[0], d[1], d[2]), r.value

    def getSortOrder(self):
        n = self.getParticleCount()
        cells = np.zeros(n, np.uint32)
        idx = np.zeros(n, np.uint32)
        check(lib().cf_get_sort_order(self._h, _ip(cells, C.c_uint32), _ip(idx, C.c_uint32), n))
        return cells, idx

    def profileSteps(self, params: SimulationParams, steps: int, edges=None):
        """Average per-kernel durations (ms) from HIP events on the engine's stream."""
        p = params.to_c()
        eo = None if edges is None else _abi.cf_edge_opts(float(edges[0]), int(edges[1]))
        ms = (C.c_double * 16)()
        nk = C.c_int()
        names = C.create_string_buffer(512)
        check(lib().cf_profile_steps(self._h, C.byref(p), int(steps),
                                     None if eo is None else C.byref(eo), ms, 16, C.byref(nk),
                                     names, 512))
        keys = names.value.decode().split(",")
        return {k: ms[i] for i, k in enumerate(keys[: nk.value])}

    def setTuning(self, cell_divisor=1, force_variant=0, lanes_per_particle=16, block_size=256,
                  steps_per_graph=32, sort_variant=1):
        t = _abi.cf_tuning(cell_divisor, force_variant, lanes_per_particle, block_size,
                           steps_per_graph, sort_variant)
        check(lib().cf_set_tuning(self._h, C.byref(t)))

    def getTuning(self):
        t = _abi.cf_tuning()
        check(lib().cf_get_tuning(self._h, C.byref(t)))
        return _tuning_dict(t)

    def autotune(self, params: SimulationParams, edges=None, trial_steps: int = 32, cache=None,
                 force: bool = False) -> dict:
        """Measure every force-kernel tuning on a copy of the current state and apply the
        fastest (the state is left unchanged).  `cache` is a JSON file path: a hit skips the
        measurement, a miss is stored.  Returns {"tuning", "us_per_step", "key", "cached",
        "trials"}."""
        p = params.to_c()
        eo = None if edges is None else _abi.cf_edge_opts(float(edges[0]), int(edges[1]))
        t = _abi.cf_tuning()
        us = C.c_double()
        rep = C.create_string_buffer(16384)
        check(lib().cf_autotune(self._h, C.byref(p), None if eo is None else C.byref(eo),
                                int(trial_steps), _path(cache),
                                _abi.CF_TUNE_FORCE if force else 0, C.byref(t), C.byref(us),
                                rep, len(rep)))
        return _tune_result(t, us, rep)

    def tuningKey(self, params: SimulationParams, edges=None) -> str:
        """The workload bucket key under which :meth:`autotune` caches its result."""
        p = params.to_c()
        eo = None if edges is None else _abi.cf_edge_opts(float(edges[0]), int(edges[1]))
        buf = C.create_string_buffer(256)
        check(lib().cf_tuning_key(self._h, C.byref(p), None if eo is None else C.byref(eo), buf,
                                  len(buf)))
        return buf.value.decode()


class Ensemble:
    """Independent simulations stepped together by the engine's kernels (no reference
    counterpart).  sims: list of (physics dict, state dict) as made by
    :func:`cellflow_amd.configs.make_ensemble`."""

    def __init__(self, sims):
        counts = (C.c_int * len(sims))(*[len(st["type"]) for _, st in sims])
        types = (C.c_int * len(sims))(*[ph["num_types"] for ph, _ in sims])
        h = C.c_void_p()
        check(lib().cf_ensemble_create(len(sims), counts, types, C.byref(h)))
        self._h = h
        self.n = [len(st["type"]) for _, st in sims]
        for s, (ph, st) in enumerate(sims):
            self.set_sim(s, ph, st)

    def set_sim(self, s, phys, st):
        p = SimulationParams.from_physics(phys).to_c()
        table = np.ascontiguousarray(st["table"], np.float32).ravel()
        rbt = np.ascontiguousarray(st["rbt"], np.float32)
        parts = particles_from_state(st["pos"], st["vel"], st["type"])
        check(lib().cf_ensemble_set_sim(self._h, s, C.byref(p), _fp(table), _fp(rbt),
                                        parts.ctypes.data_as(C.POINTER(_abi.cf_particle)),
                                        len(parts)))

    def simulate(self, steps: int):
        check(lib().cf_ensemble_simulate(self._h, int(steps)))

    def getParticleData(self, s: int) -> np.ndarray:
        out = np.zeros(self.n[s], PARTICLE_DTYPE)
        check(lib().cf_ensemble_get_particles(self._h, s,
                                              out.ctypes.data_as(C.POINTER(_abi.cf_particle)),
                                              self.n[s]))
        return out

    def getNeighborCounts(self, s: int) -> np.ndarray:
        c = np.zeros(self.n[s], np.int32)
        check(lib().cf_ensemble_get_neighbor_counts(self._h, s, _ip(c), self.n[s]))
        return c

    def setTuning(self, cell_divisor=1, force_variant=0, lanes_per_particle=16, block_size=256,
                  steps_per_graph=32, sort_variant=1):
        t = _abi.cf_tuning(cell_divisor, force_variant, lanes_per_particle, block_size,
                           steps_per_graph, sort_variant)
        check(lib().cf_ensemble_set_tuning(self._h, C.byref(t)))

    def getTuning(self):
        t = _abi.cf_tuning()
        check(lib().cf_ensemble_get_tuning(self._h, C.byref(t)))
        return _tuning_dict(t)

    def autotune(self, trial_steps: int = 32, cache=None, force: bool = False) -> dict:
        """:meth:`ParticleSimulation.autotune` for the whole ensemble."""
        t = _abi.cf_tuning()
        us = C.c_double()
        rep = C.create_string_buffer(16384)
        check(lib().cf_ensemble_autotune(self._h, int(trial_steps), _path(cache),
                                         _abi.CF_TUNE_FORCE if force else 0, C.byref(t),
                                         C.byref(us), rep, len(rep)))
        return _tune_result(t, us, rep)

    def close(self):
        if self._h:
            check(lib().cf_ensemble_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().cf_ensemble_destroy(self._h)
        except Exception:
            pass


def _tuning_dict(t) -> dict:
    return dict(cell_divisor=t.cell_divisor, force_variant=t.force_variant,
                lanes_per_particle=t.lanes_per_particle, block_size=t.block_size,
                steps_per_graph=t.steps_per_graph, sort_variant=t.sort_variant)


def _path(p):
    return None if p is None else os.fsencode(p)


def _tune_result(t, us, rep) -> dict:
    r = json.loads(rep.value.decode())
    return dict(tuning=_tuning_dict(t), us_per_step=us.value, key=r["key"], cached=r["cached"],
                trials=r.get("trials", []))


def tuning_cache_lookup(path, key: str):
    """(tuning dict, us_per_step) stored under `key` in the cache file, or None."""
    t = _abi.cf_tuning()
    us = C.c_double()
    found = C.c_int()
    check(lib().cf_tuning_cache_lookup(os.fsencode(path), key.encode(), C.byref(t), C.byref(us),
                                       C.byref(found)))
    return (_tuning_dict(t), us.value) if found.value else None


def tuning_cache_store(path, key: str, tuning: dict, us_per_step: float) -> None:
    t = _abi.cf_tuning(tuning["cell_divisor"], tuning["force_variant"],
                       tuning["lanes_per_particle"], tuning["block_size"],
                       tuning["steps_per_graph"], tuning.get("sort_variant", 1))
    check(lib().cf_tuning_cache_store(os.fsencode(path), key.encode(), C.byref(t),
                                      float(us_per_step)))


def particles_from_state(pos, vel, type_) -> np.ndarray:
    n = len(type_)
    a = np.zeros(n, PARTICLE_DTYPE)
    a["pos"] = pos
    a["vel"] = vel
    a["ptype"] = type_
    return a


def grid_dims(params: SimulationParams, rbt, n: int, cell_divisor: int = 1):
    p = params.to_c()
    r = np.ascontiguousarray(rbt, np.float32)
    d = (C.c_int * 3)()
    check(lib().cf_grid_dims(C.byref(p), _fp(r), int(n), int(cell_divisor), d))
    return (d[0], d[1], d[2])


def cutoff_threshold(eff_radius: float) -> float:
    return float(lib().cf_cutoff_threshold(float(eff_radius)))


def force_table_transform(raw, rng, bias, offset) -> np.ndarray:
    r = np.ascontiguousarray(raw, np.float32).ravel()
    out = np.zeros_like(r)
    lib().cf_force_table_transform(_fp(r), r.size, rng, bias, offset, _fp(out))
    return out


def device_count() -> int:
    n = C.c_int()
    rc = lib().cf_device_count(C.byref(n))
    return n.value if rc == 0 else 0
