"""GPU parity of the triangle mesh (generateTriangleMesh, ParticleSimulation.cu:689-754,
and the triangle-vertex layout of ParticleSimulation.cu:362-418) against the oracle.

Triangles are index work and must match the oracle's canonical list (owner ascending, then
fan order) bit for bit.  Vertex positions and colours are copies and must be exact; the
flat normals are the same FP32 expression in both and are held to 1e-6 absolute.
"""

import numpy as np
import pytest

from cellflow_amd import CellFlowError, ParticleSimulation, _abi
from helpers import make_sim, small_cfg, state_of

pytestmark = pytest.mark.gpu

NORMAL_ATOL = 1e-6


def clustered(n=5000, side=1.0, T=2, blobs=4, sd=0.006, seed=3):
    cfg = small_cfg(n=n, side=side, T=T, seed=seed)
    cfg["init"] = dict(kind="blobs", count=blobs, sd=sd)
    return state_of(cfg)


def check_vertices(orc, sim, tris, T):
    got = sim.getParticleData()
    rng = np.random.default_rng(T)
    colors = rng.random((T, 3)).astype(np.float32)
    v = sim.getTriangleVertices(colors)
    ref = orc.triangle_vertices(tris, got["pos"], got["ptype"], T, colors)
    assert v.shape == (3 * len(tris), 9)
    assert np.array_equal(v[:, :3], ref[:, :3])
    assert np.array_equal(v[:, 6:], ref[:, 6:])
    np.testing.assert_allclose(v[:, 3:6], ref[:, 3:6], rtol=0, atol=NORMAL_ATOL)


@pytest.mark.parametrize("maxc", [1, 4, 16, 32, 40])
def test_mesh_dense_blobs_matches_oracle(orc, maxc):
    """Dense blobs: most owners have far more same-type neighbours than the list holds, so
    the index-order truncation and the 64-triangle cap both matter (maxConn 1 keeps one
    neighbour per owner, hence no triangle)."""
    st, phys, params = clustered()
    sim = make_sim(st, phys)
    sim.simulate(params, 0)  # the blobs as seeded (a step spreads them)
    got = sim.getParticleData()
    for dist in (0.004, 0.012, 0.15):  # 0.15 * 1.2 spans several cells
        t = sim.generateTriangleMesh(dist, maxc)
        ref = orc.triangles(got["pos"], got["ptype"], dist, maxc, mode=1)
        assert np.array_equal(t, ref), f"dist={dist} maxc={maxc}: {len(t)} vs {len(ref)}"
        assert np.array_equal(sim.getTriangles(), t)
        assert (len(t) > 0) == (maxc > 1)
        if dist == 0.004:
            check_vertices(orc, sim, t, 2)


@pytest.mark.parametrize("name,over", [
    ("uniform", dict(n=4000, side=0.5, T=6)),
    ("ten_types", dict(n=3000, side=0.4, T=10)),
    ("tiny_grid", dict(n=400, side=0.08, T=3)),
])
def test_mesh_uniform_matches_oracle(orc, name, over):
    st, phys, params = state_of(small_cfg(**over))
    sim = make_sim(st, phys)
    sim.simulate(params, 2, edges=(0.03, 5))  # a graph beside the mesh changes neither
    got = sim.getParticleData()
    t = sim.generateTriangleMesh(0.03, 8)
    ref = orc.triangles(got["pos"], got["ptype"], 0.03, 8, mode=1)
    assert np.array_equal(t, ref)
    assert len(t) > 0
    assert np.array_equal(sim.getEdges(), orc.edges(got["pos"], got["ptype"], 0.03, 5, mode=0))
    check_vertices(orc, sim, t, over["T"])


def test_mesh_non_cubic_canvas(orc):
    st, phys, params = state_of(small_cfg(n=3000, side=0.5, canvas=[0.8, 0.3, 0.45]))
    sim = make_sim(st, phys)
    sim.simulate(params, 1)
    got = sim.getParticleData()
    t = sim.generateTriangleMesh(0.04, 6)
    assert np.array_equal(t, orc.triangles(got["pos"], got["ptype"], 0.04, 6, mode=1))


def test_mesh_rebuild_shrinks_and_zero_distance(orc):
    st, phys, params = state_of(small_cfg(n=3000, side=0.5))
    sim = make_sim(st, phys)
    sim.simulate(params, 0)
    big = sim.generateTriangleMesh(0.05, 12)
    small = sim.generateTriangleMesh(0.02, 4)
    assert 0 < len(small) < len(big)
    assert np.array_equal(small, orc.triangles(st["pos"], st["type"], 0.02, 4, mode=1))
    assert len(sim.generateTriangleMesh(0.0, 8)) == 0
    assert len(sim.generateTriangleMesh(0.05, 0)) == 0
    assert sim.getTriangleVertices(np.ones((6, 3), np.float32)).shape == (0, 9)


def test_mesh_errors():
    sim = ParticleSimulation(100, 2)
    with pytest.raises(CellFlowError):
        sim.generateTriangleMesh(0.02, 4)  # no parameters yet
    with pytest.raises(CellFlowError):
        sim.getTriangles()  # nothing built yet
    st, phys, params = state_of(small_cfg(n=500, side=0.3, T=2))
    sim = make_sim(st, phys)
    sim.simulate(params, 0)
    with pytest.raises(CellFlowError):
        sim.generateTriangleMesh(float("nan"), 4)
    with pytest.raises(CellFlowError):
        sim.generateTriangleMesh(-1.0, 4)


def test_vertex_buffers_reused_and_mesh_dropped_with_a_new_state(orc):
    """The vertex expansions share device buffers across calls (they grow on demand); a
    state loaded after a build drops the graph and the mesh, which belong to the old one."""
    st, phys, params = state_of(small_cfg(n=3000, side=0.5))
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.03, 4))
    got = sim.getParticleData()
    small = sim.generateTriangleMesh(0.02, 4)
    check_vertices(orc, sim, small, 6)
    big = sim.generateTriangleMesh(0.05, 12)
    assert len(big) > 4 * len(small) > 0
    check_vertices(orc, sim, big, 6)
    colors = np.eye(3, dtype=np.float32)[np.arange(6) % 3]
    edges = sim.getEdges()
    lv = sim.getLineVertices(colors)
    assert np.array_equal(lv[0::2, :3], got["pos"][edges[:, 0]])
    assert np.array_equal(lv[1::2, :3], got["pos"][edges[:, 1]])
    assert np.array_equal(lv[0::2, 3:], colors[got["ptype"][edges[:, 0]] % 6])
    check_vertices(orc, sim, big, 6)
    sim.setParticles(got)
    for read in (sim.getTriangles, sim.getEdges, lambda: sim.getTriangleVertices(colors),
                 lambda: sim.getLineVertices(colors)):
        with pytest.raises(CellFlowError):
            read()


def test_mesh_dropped_when_the_state_advances():
    """A mesh belongs to the state it was built on: a step, a profiled step or a gradient
    rollout advances the state, after which reading the mesh is CF_ERR_STATE.  A call that
    does not step (simulate with 0 steps) keeps it."""
    st, phys, params = state_of(small_cfg(n=2000, side=0.4))
    sim = make_sim(st, phys)
    sim.simulate(params, 1)
    colors = np.ones((6, 3), np.float32)
    advance = [lambda: sim.simulate(params, 1),
               lambda: sim.profileSteps(params, 2),
               lambda: sim.rolloutGrad(params, 2, 1, 8, 0.1, np.zeros(8))]
    for step in advance:
        tris = sim.generateTriangleMesh(0.03, 8)
        assert len(tris) > 0
        sim.simulate(params, 0)
        assert np.array_equal(sim.getTriangles(), tris)
        step()
        for read in (sim.getTriangles, lambda: sim.getTriangleVertices(colors)):
            with pytest.raises(CellFlowError) as e:
                read()
            assert e.value.code == _abi.CF_ERR_STATE
    sim.close()


def _hip():
    """The HIP runtime the engine library runs on, through ctypes (dlopen by its soname
    returns the copy already loaded), for raw allocations in tests.  Torch's own device
    memory is not used here: torch may bring a second copy of the runtime into the process."""
    import ctypes as C
    from cellflow_amd import _abi
    _abi.lib()  # the engine and its runtime are loaded first
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipHostFree.argtypes = [C.c_void_p]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipPointerGetAttributes.argtypes = [C.POINTER(_PtrAttr), C.c_void_p]
    return h


class _PtrAttr(__import__("ctypes").Structure):
    import ctypes as C
    _fields_ = [("type", C.c_int), ("device", C.c_int), ("devicePointer", C.c_void_p),
                ("hostPointer", C.c_void_p), ("isManaged", C.c_int),
                ("allocationFlags", C.c_uint)]


def test_vertices_written_into_device_memory(orc):
    """The device-side vertex writes (the reference maps its GL buffer and the kernels write
    into it) give the same bytes as the host expansions, into device memory and into pinned
    host memory; too small an output is CF_ERR_CAPACITY, both by the stated capacity and by
    the allocation behind the pointer, with nothing launched."""
    import ctypes as C
    st, phys, params = state_of(small_cfg(n=3000, side=0.5))
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.03, 6))
    colors = np.random.default_rng(0).random((6, 3)).astype(np.float32)
    lines = sim.getLineVertices(colors)
    sim.generateTriangleMesh(0.03, 8)
    tverts = sim.getTriangleVertices(colors)
    assert len(lines) > 100 and len(tverts) > 100
    hip = _hip()
    for ref, write, width in ((lines, sim.writeLineVerticesDevice, 6),
                              (tverts, sim.writeTriangleVerticesDevice, 9)):
        need = ref.size
        dev, pin = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(dev), 4 * (need + 64)) == 0
        assert hip.hipHostMalloc(C.byref(pin), 4 * need, 0) == 0
        try:
            assert hip.hipMemset(dev, 0x7F, 4 * (need + 64)) == 0
            assert write(colors, dev.value, need + 64) == len(ref)
            back = np.zeros(need + 64, np.float32)
            assert hip.hipMemcpy(back.ctypes.data, dev, 4 * (need + 64), 2) == 0  # to host
            assert back[:need].tobytes() == ref.tobytes()
            assert back[need:].view(np.uint32).tolist() == [0x7F7F7F7F] * 64  # nothing past
            assert write(colors, pin.value, need) == len(ref)
            pinned = np.ctypeslib.as_array(C.cast(pin, C.POINTER(C.c_float)), (need,))
            assert pinned.tobytes() == ref.tobytes()
            with pytest.raises(CellFlowError) as e:
                write(colors, dev.value, need - 1)
            assert e.value.code == _abi.CF_ERR_CAPACITY
            # a capacity beyond the allocation behind the pointer is caught too
            with pytest.raises(CellFlowError) as e:
                write(colors, dev.value + 4 * (64 + width), need)
            assert e.value.code == _abi.CF_ERR_CAPACITY
        finally:
            hip.hipFree(dev)
            hip.hipHostFree(pin)
    sim.close()


def test_device_vertex_write_rejects_pageable_host_memory():
    """Pageable host memory is not writable by the GPU here: the call must refuse it on the
    host.  The case runs only when the runtime itself reports the buffer as unregistered or
    unknown, which is what the engine's check tests before launching anything."""
    import ctypes as C
    st, phys, params = state_of(small_cfg(n=2000, side=0.4))
    sim = make_sim(st, phys)
    sim.simulate(params, 1, edges=(0.03, 6))
    colors = np.ones((6, 3), np.float32)
    need = sim.getLineVertices(colors).size
    host = np.zeros(need, np.float32)
    hip = _hip()
    at = _PtrAttr()
    rc = hip.hipPointerGetAttributes(C.byref(at), host.ctypes.data)
    hip.hipGetLastError()  # leave no error from the query for later calls to find
    if rc == 0 and at.type != 0:
        pytest.skip(f"the runtime reports pageable memory as type {at.type}")
    with pytest.raises(CellFlowError) as e:
        sim.writeLineVerticesDevice(colors, host.ctypes.data, need)
    assert e.value.code == _abi.CF_ERR_ARG
    assert not host.any()
    sim.close()
