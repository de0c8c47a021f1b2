"""CPU oracle for the CellFlow step — TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/_build/liboracle.so`` (built from ``oracle/oracle.cpp`` by
``oracle/Makefile``).  Only the tests, the smoke check and the benchmark's CPU-baseline
leg import this module, and only to compare the engine against.  The engine never calls it.

Each wrapper names the reference lines its C function restates (paths relative to the
reference root):

* :func:`step`          cuda-native/src/ParticleSimulation.cu:69-166 (Jacobi form)
* :func:`rollout`       ParticleSimulation.cu:541-556 (count ping-pong over steps)
* :func:`edges`         ParticleSimulation.cu:187-246
* :func:`triangles`     ParticleSimulation.cu:279-348 with the driver's threshold :722-724
* :func:`triangle_vertices`  ParticleSimulation.cu:350-418
* :func:`force_table`   ParticleSimulation.cu:521-531
* :func:`move`          ParticleSimulation.cu:169-185
* :func:`bin_cells`, :func:`order`   the engine's grid rule (no reference counterpart)
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


class OrcParams(C.Structure):
    _fields_ = [
        ("radius", C.c_float), ("dt", C.c_float), ("friction", C.c_float),
        ("repulsion", C.c_float), ("attraction", C.c_float), ("k", C.c_float),
        ("balance", C.c_float), ("L", C.c_float * 3), ("T", C.c_int), ("ratio", C.c_float),
        ("force_multiplier", C.c_float), ("max_expected_neighbors", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ only; no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-C", _HERE, "-B"] if force else ["make", "-C", _HERE]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.orc_step.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_uint32), P(C.c_int32),
                               P(OrcParams), P(C.c_float), P(C.c_float), P(C.c_double),
                               P(C.c_double), P(C.c_double), P(C.c_int32), P(C.c_double),
                               C.c_int, C.c_int]
        L.orc_rollout.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_uint32),
                                  P(C.c_int32), P(OrcParams), P(C.c_float), P(C.c_float),
                                  C.c_int, C.c_int, C.c_int]
        L.orc_bin.argtypes = [C.c_int, P(C.c_float), P(C.c_int), P(C.c_float), P(C.c_uint32)]
        L.orc_order.argtypes = [C.c_int, P(C.c_uint32), P(C.c_uint32)]
        L.orc_edges.argtypes = [C.c_int, P(C.c_float), P(C.c_uint32), C.c_float, C.c_int, C.c_int,
                                P(C.c_int32), C.c_longlong, P(C.c_longlong), C.c_int]
        L.orc_triangles.argtypes = [C.c_int, P(C.c_float), P(C.c_uint32), C.c_float, C.c_int,
                                    C.c_int, P(C.c_int32), C.c_longlong, P(C.c_longlong), C.c_int]
        L.orc_triangle_vertices.argtypes = [C.c_longlong, P(C.c_int32), P(C.c_float),
                                            P(C.c_uint32), C.c_int, P(C.c_float), C.c_int,
                                            P(C.c_float)]
        L.orc_triangle_vertices.restype = None
        L.orc_force_table.argtypes = [P(C.c_float), C.c_int, C.c_float, C.c_float, C.c_float,
                                      P(C.c_float)]
        L.orc_force_table.restype = None
        L.orc_move.argtypes = [C.c_int, P(C.c_float), C.c_float, C.c_float, C.c_float,
                               P(C.c_float)]
        L.orc_move.restype = None
        L.orc_pair_in_range.argtypes = [P(OrcParams), P(C.c_float), P(C.c_float), P(C.c_float),
                                        C.c_uint, C.c_uint]
        L.orc_eff_radius.argtypes = [P(OrcParams), P(C.c_float), C.c_uint, C.c_uint]
        L.orc_eff_radius.restype = C.c_float
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def make_params(phys: dict) -> OrcParams:
    p = OrcParams()
    p.radius = phys["radius"]
    p.dt = phys["delta_t"]
    p.friction = phys["friction"]
    p.repulsion = phys["repulsion"]
    p.attraction = phys["attraction"]
    p.k = phys["k"]
    p.balance = phys["balance"]
    p.L[0], p.L[1], p.L[2] = phys["canvas"]
    p.T = phys["num_types"]
    p.ratio = phys["ratio"]
    p.force_multiplier = phys["force_multiplier"]
    p.max_expected_neighbors = phys["max_expected_neighbors"]
    return p


def step(pos, vel, ptype, prev, phys, table, rbt, mode=1, nthreads=0):
    """One Jacobi step (FP64 continuous math, FP32 decisions).

    Returns dict with pos, vel, acc (float64 N x 3), count (int32) and fabs (sum of |net|)."""
    n = len(ptype)
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    ptype = np.ascontiguousarray(ptype, np.uint32)
    prev = np.ascontiguousarray(prev, np.int32)
    table = np.ascontiguousarray(table, np.float32).ravel()
    rbt = np.ascontiguousarray(rbt, np.float32)
    out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "acc")}
    out["count"] = np.zeros(n, np.int32)
    out["fabs"] = np.zeros(n)
    p = make_params(phys)
    rc = lib().orc_step(n, _p(pos, C.c_double), _p(vel, C.c_double), _p(ptype, C.c_uint32),
                        _p(prev, C.c_int32), C.byref(p), _p(table, C.c_float), _p(rbt, C.c_float),
                        _p(out["pos"], C.c_double), _p(out["vel"], C.c_double),
                        _p(out["acc"], C.c_double), _p(out["count"], C.c_int32),
                        _p(out["fabs"], C.c_double), mode, nthreads)
    if rc:
        raise ValueError(f"orc_step failed rc={rc}")
    return out


def rollout(pos, vel, ptype, prev, phys, table, rbt, steps, mode=1, nthreads=0):
    """`steps` Jacobi steps in FP64 (FP32 decisions). Returns (pos, vel, prev) copies."""
    pos = np.array(pos, np.float64, order="C")
    vel = np.array(vel, np.float64, order="C")
    prev = np.array(prev, np.int32, order="C")
    ptype = np.ascontiguousarray(ptype, np.uint32)
    table = np.ascontiguousarray(table, np.float32).ravel()
    rbt = np.ascontiguousarray(rbt, np.float32)
    p = make_params(phys)
    rc = lib().orc_rollout(len(ptype), _p(pos, C.c_double), _p(vel, C.c_double),
                           _p(ptype, C.c_uint32), _p(prev, C.c_int32), C.byref(p),
                           _p(table, C.c_float), _p(rbt, C.c_float), steps, mode, nthreads)
    if rc:
        raise ValueError(f"orc_rollout failed rc={rc}")
    return pos, vel, prev


def bin_cells(pos, dims, L):
    pos = np.ascontiguousarray(pos, np.float32)
    dims = np.ascontiguousarray(dims, np.int32)
    L = np.ascontiguousarray(L, np.float32)
    out = np.zeros(len(pos), np.uint32)
    rc = lib().orc_bin(len(pos), _p(pos, C.c_float), _p(dims, C.c_int), _p(L, C.c_float),
                       _p(out, C.c_uint32))
    if rc:
        raise ValueError("orc_bin failed")
    return out


def order(cells):
    cells = np.ascontiguousarray(cells, np.uint32)
    out = np.zeros(len(cells), np.uint32)
    lib().orc_order(len(cells), _p(cells, C.c_uint32), _p(out, C.c_uint32))
    return out


def edges(pos, ptype, dist, max_conn, mode=1, nthreads=0):
    """Canonical proximity-graph pairs, shape (E, 2) int32, owner ascending then rank."""
    pos = np.ascontiguousarray(pos, np.float32)
    ptype = np.ascontiguousarray(ptype, np.uint32)
    n = len(ptype)
    cap = max(1, n * max(0, min(max_conn, 32)))
    out = np.zeros((cap, 2), np.int32)
    cnt = C.c_longlong(0)
    rc = lib().orc_edges(n, _p(pos, C.c_float), _p(ptype, C.c_uint32), float(dist), int(max_conn),
                         mode, _p(out, C.c_int32), cap, C.byref(cnt), nthreads)
    if rc:
        raise ValueError(f"orc_edges failed rc={rc}")
    return out[: cnt.value].copy()


def triangles(pos, ptype, dist, max_conn, mode=1, nthreads=0):
    """Canonical triangle mesh, shape (T, 3) int32: owner ascending, then fan order."""
    pos = np.ascontiguousarray(pos, np.float32)
    ptype = np.ascontiguousarray(ptype, np.uint32)
    n = len(ptype)
    cap = max(1, n * 64)
    out = np.zeros((cap, 3), np.int32)
    cnt = C.c_longlong(0)
    rc = lib().orc_triangles(n, _p(pos, C.c_float), _p(ptype, C.c_uint32), float(dist),
                             int(max_conn), mode, _p(out, C.c_int32), cap, C.byref(cnt), nthreads)
    if rc:
        raise ValueError(f"orc_triangles failed rc={rc}")
    return out[: cnt.value].copy()


def triangle_vertices(tris, pos, ptype, num_types, colors):
    """(3T, 9) float32 vertex rows (position, flat normal, owner colour) of `tris`."""
    tris = np.ascontiguousarray(tris, np.int32)
    pos = np.ascontiguousarray(pos, np.float32)
    ptype = np.ascontiguousarray(ptype, np.uint32)
    col = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
    out = np.zeros((max(len(tris), 1) * 3, 9), np.float32)
    lib().orc_triangle_vertices(len(tris), _p(tris, C.c_int32), _p(pos, C.c_float),
                                _p(ptype, C.c_uint32), int(num_types), _p(col, C.c_float),
                                len(col), _p(out, C.c_float))
    return out[: 3 * len(tris)]


def force_table(raw, rng, bias, offset):
    raw = np.ascontiguousarray(raw, np.float32).ravel()
    out = np.zeros_like(raw)
    lib().orc_force_table(_p(raw, C.c_float), len(raw), rng, bias, offset, _p(out, C.c_float))
    return out


def move(pos, dx, dy, dz, L):
    """moveUniverse's wrap (ParticleSimulation.cu:169-185) on a copy of FP32 positions."""
    out = np.array(pos, np.float32, order="C")
    Lf = np.ascontiguousarray(L, np.float32)
    lib().orc_move(len(out), _p(out, C.c_float), dx, dy, dz, _p(Lf, C.c_float))
    return out


def pair_in_range(phys, rbt, xi, xj, ti, tj):
    p = make_params(phys)
    rbt = np.ascontiguousarray(rbt, np.float32)
    xi = np.ascontiguousarray(xi, np.float32)
    xj = np.ascontiguousarray(xj, np.float32)
    return bool(lib().orc_pair_in_range(C.byref(p), _p(rbt, C.c_float), _p(xi, C.c_float),
                                         _p(xj, C.c_float), ti, tj))


def eff_radius(phys, rbt, ti, tj):
    p = make_params(phys)
    rbt = np.ascontiguousarray(rbt, np.float32)
    return lib().orc_eff_radius(C.byref(p), _p(rbt, C.c_float), ti, tj)


# ---- differentiable rollout ---------------------------------------------------------------

def _bind_adjoint(L):
    P = C.POINTER
    if getattr(L, "_adj_bound", False):
        return L
    L.orc_hist.argtypes = [C.c_int, P(C.c_double), P(C.c_float), C.c_int, C.c_double,
                           P(C.c_double), C.c_int]
    L.orc_pairs.argtypes = [C.c_int, P(C.c_float), P(C.c_uint32), P(OrcParams), P(C.c_float),
                            P(C.c_int64), P(C.c_int32), C.c_int64, C.c_int]
    L.orc_frozen_rollout.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_uint32),
                                     P(C.c_int32), P(OrcParams), P(C.c_float), P(C.c_double),
                                     C.c_double, P(C.c_float), C.c_int, P(P(C.c_int64)),
                                     P(P(C.c_int32)), C.c_int, C.c_double, P(C.c_double),
                                     P(C.c_double), P(C.c_double), P(C.c_double), C.c_int]
    L._adj_bound = True
    return L


def hist(pos, L, bins, r_max, nthreads=0):
    """Soft pair histogram h_b = (1/N) sum_{i != j} B3((r_ij - c_b)/w) of FP64 positions."""
    pos = np.ascontiguousarray(pos, np.float64)
    Lf = np.ascontiguousarray(L, np.float32)
    out = np.zeros(bins)
    r_max = float(np.float32(r_max))  # the engine's ABI takes r_max as FP32
    _bind_adjoint(lib()).orc_hist(len(pos), _p(pos, C.c_double), _p(Lf, C.c_float), bins,
                                  r_max, _p(out, C.c_double), nthreads)
    return out


def pairs(pos, ptype, phys, rbt, nthreads=0):
    """Neighbour lists (CSR offsets, indices) of one state by the FP32 predicate."""
    pos = np.ascontiguousarray(pos, np.float32)
    ptype = np.ascontiguousarray(ptype, np.uint32)
    rbt = np.ascontiguousarray(rbt, np.float32)
    n = len(ptype)
    p = make_params(phys)
    cap = n * 64
    while True:
        offs = np.zeros(n + 1, np.int64)
        nbr = np.zeros(max(cap, 1), np.int32)
        rc = _bind_adjoint(lib()).orc_pairs(n, _p(pos, C.c_float), _p(ptype, C.c_uint32),
                                            C.byref(p), _p(rbt, C.c_float), _p(offs, C.c_int64),
                                            _p(nbr, C.c_int32), cap, nthreads)
        if rc == 0:
            return offs, nbr[: offs[-1]].copy()
        cap *= 4


class FrozenRollout:
    """FP64 rollouts whose per-step neighbour sets (and counts) are fixed in advance, e.g. to
    the decisions of a given trajectory.  loss(table, ratio) is then a smooth function of the
    parameters, which central finite differences can differentiate."""

    def __init__(self, pos0, vel0, ptype, prev0, phys, rbt, lists, bins, r_max, target,
                 nthreads=0):
        self.pos0 = np.ascontiguousarray(pos0, np.float64)
        self.vel0 = np.ascontiguousarray(vel0, np.float64)
        self.ptype = np.ascontiguousarray(ptype, np.uint32)
        self.prev0 = np.ascontiguousarray(prev0, np.int32)
        self.p = make_params(phys)
        self.rbt = np.ascontiguousarray(rbt, np.float32)
        self.lists = lists
        self.steps = len(lists)
        self._offs = (C.POINTER(C.c_int64) * self.steps)(*[_p(o, C.c_int64) for o, _ in lists])
        self._nbr = (C.POINTER(C.c_int32) * self.steps)(*[_p(b, C.c_int32) for _, b in lists])
        self.bins, self.r_max = bins, float(np.float32(r_max))  # FP32 like the engine's ABI
        self.target = np.ascontiguousarray(target, np.float64)
        self.nthreads = nthreads

    def loss(self, table, ratio, want_hist=False):
        table = np.ascontiguousarray(table, np.float64).ravel()
        tf = table.astype(np.float32)
        out = C.c_double()
        h = np.zeros(self.bins)
        _bind_adjoint(lib()).orc_frozen_rollout(
            len(self.ptype), _p(self.pos0, C.c_double), _p(self.vel0, C.c_double),
            _p(self.ptype, C.c_uint32), _p(self.prev0, C.c_int32), C.byref(self.p),
            _p(tf, C.c_float), _p(table, C.c_double), float(ratio), _p(self.rbt, C.c_float),
            self.steps, self._offs, self._nbr, self.bins, self.r_max,
            _p(self.target, C.c_double), C.byref(out), None, _p(h, C.c_double), self.nthreads)
        return (out.value, h) if want_hist else out.value

    def fd_gradient(self, table, ratio, rel_step=1e-4, abs_step=1e-6):
        """Central differences with respect to the T*T table entries and the ratio."""
        table = np.asarray(table, np.float64)
        T = table.shape[0]
        g = np.zeros(T * T + 1)
        for k in range(T * T + 1):
            if k < T * T:
                x0 = table.flat[k]
            else:
                x0 = ratio
            h = max(abs(x0) * rel_step, abs_step)
            vals = []
            for sgn in (1, -1):
                tb = table.copy()
                rt = ratio
                if k < T * T:
                    tb.flat[k] = x0 + sgn * h
                else:
                    rt = x0 + sgn * h
                vals.append(self.loss(tb, rt))
            g[k] = (vals[0] - vals[1]) / (2 * h)
        return g[: T * T].reshape(T, T), g[T * T]
