"""The header-only C++ drop-in (include/cellflow.hpp) through tests/cpp/wrapper_check.cpp,
built by __graft_entry__.build().

Without a device the program must fail cleanly with the device status.  On a GPU it drives
the wrapper the way the reference's widget drives its class (one step per frame, proximity
graph after each step, raw-table and radius edits); replaying the same calls through the
Python binding must give the same bytes."""

import os
import subprocess

import numpy as np
import pytest

from cellflow_amd import PARTICLE_DTYPE, ParticleSimulation, SimulationParams, device_count
from cellflow_amd import force_table_transform, read_preset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "wrapper_check")


PRESET = os.path.join(ROOT, "tests", "golden", "presets", "eater.json")


def run(outdir):
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    return subprocess.run([BIN, str(outdir), PRESET], capture_output=True, text=True,
                          timeout=120)


@pytest.mark.skipif(device_count() > 0, reason="checks the no-device path")
def test_wrapper_without_device_reports_status(tmp_path):
    r = run(tmp_path)
    assert r.returncode == 12, r.stdout + r.stderr
    assert "no HIP device" in r.stdout


@pytest.mark.gpu
def test_wrapper_autotune():
    """The class's autotune(): the state is unchanged, the chosen tuning is applied, the
    tuned simulation steps."""
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([BIN, "--tune"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tuned: variant" in r.stdout
    print(r.stdout.strip())


@pytest.mark.gpu
def test_wrapper_matches_python_binding(tmp_path):
    r = run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip())

    def load(name, dtype):
        return np.fromfile(tmp_path / name, dtype=dtype)

    init = load("init.bin", PARTICLE_DTYPE)
    raw = load("raw.bin", np.float32)
    table = load("table.bin", np.float32)
    rbt = load("rbt.bin", np.float32)
    # the wrapper's table is updateForceTable(0.6, 0.1, 1.5) of its raw values
    np.testing.assert_array_equal(table, force_table_transform(raw, 0.6, 0.1, 1.5))
    assert raw[3] == np.float32(0.5) and np.all(np.abs(raw) <= 1)
    assert np.all(init["pos"] >= 0) and np.all(init["pos"] < np.float32(0.6))
    assert not init["vel"].any()

    sim = ParticleSimulation(len(init), 6)
    sim.setRawForceTableValues(raw)
    sim.updateForceTable(0.6, 0.1, 1.5)
    np.testing.assert_array_equal(sim.getForceTable().ravel(), table)
    sim.setRadioByType(rbt)
    sim.setParticles(init)
    p = SimulationParams(radius=0.05, delta_t=0.01, friction=0.1, k=1.0, canvasWidth=0.6,
                         canvasHeight=0.6, canvasDepth=0.6, ratioWithLFO=0.3)
    colors = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]],
                      np.float32)
    for _ in range(12):
        sim.simulate(p)
        sim.generateProximityGraph(0.02, 16)
    assert sim.getParticleData().tobytes() == load("final.bin", PARTICLE_DTYPE).tobytes()
    assert np.array_equal(sim.getNeighborCounts(), load("counts.bin", np.int32))
    assert np.array_equal(sim.getEdges().ravel(), load("edges.bin", np.int32))
    assert np.array_equal(sim.getLineVertices(colors).ravel(), load("verts.bin", np.float32))
    # the wrapper built its triangle mesh after the last graph (and read the graph after it)
    sim.generateTriangleMesh(0.05, 8)
    tverts = load("tverts.bin", np.float32)
    assert len(tverts) > 0
    assert np.array_equal(sim.getTriangleVertices(colors).ravel(), tverts)
    assert np.array_equal(sim.getEdges().ravel(), load("edges.bin", np.int32))
    # the preset loaded, stepped and saved by the C++ wrapper: same file, same table
    assert (tmp_path / "saved.json").read_bytes() == open(PRESET, "rb").read()
    pr = read_preset(PRESET)
    np.testing.assert_array_equal(load("preset_table.bin", np.float32),
                                  force_table_transform(pr["rawForceTable"], pr["forceRange"],
                                                        pr["forceBias"], pr["forceOffset"]))
