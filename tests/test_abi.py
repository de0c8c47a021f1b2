"""The C ABI without a GPU: the library loads, exports every symbol include/cellflow.h
declares, reports errors instead of exiting, and its host-only helpers are correct.
No call here launches a kernel."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from cellflow_amd import SimulationParams, _abi, cutoff_threshold, device_count, grid_dims

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cellflow.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|float|void|const char\*)\s+(cf_[a-z_]+)\(", text,
                                 re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _abi.lib()
    names = declared_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, f"{n} has no ctypes signature"
    assert set(_abi.SIGNATURES) == set(names)


def test_struct_layouts():
    assert C.sizeof(_abi.cf_particle) == 44
    assert _abi.cf_particle.ptype.offset == 36 and _abi.cf_particle.pad.offset == 40
    assert C.sizeof(_abi.cf_params) == 14 * 4
    assert C.sizeof(_abi.cf_tuning) == 6 * 4


def test_reference_defaults():
    """SimulationParams defaults of cuda-native/include/SimulationParams.h:15-29."""
    p = SimulationParams()
    assert (p.radius, p.delta_t, p.friction, p.repulsion, p.attraction, p.k, p.balance) == \
        (42.07, 0.18, 0.51, 64.83, 3.06, 29.45, 0.79)
    assert (p.canvasWidth, p.numParticleTypes, p.ratioWithLFO, p.forceMultiplier,
            p.maxExpectedNeighbors) == (8000.0, 6, 0.0, 2.33, 400)


@pytest.mark.skipif(device_count() > 0, reason="checks the no-device path")
def test_create_without_device_fails_cleanly():
    lib = _abi.lib()
    h = C.c_void_p()
    rc = lib.cf_create(1000, 4, C.byref(h))
    assert rc == _abi.CF_ERR_DEVICE and not h.value
    assert b"device" in lib.cf_last_error_string()
    rc = lib.cf_ensemble_create(2, (C.c_int * 2)(10, 20), (C.c_int * 2)(2, 3), C.byref(h))
    assert rc == _abi.CF_ERR_DEVICE


def test_argument_errors_before_device():
    lib = _abi.lib()
    h = C.c_void_p()
    assert lib.cf_create(0, 4, C.byref(h)) == _abi.CF_ERR_ARG
    assert lib.cf_create(100, 11, C.byref(h)) == _abi.CF_ERR_ARG
    assert lib.cf_create(100, 0, C.byref(h)) == _abi.CF_ERR_ARG
    assert lib.cf_create(100, 4, None) == _abi.CF_ERR_ARG


def test_invalid_handles_are_rejected():
    lib = _abi.lib()
    bogus = C.c_void_p(0x1234)  # never dereferenced: handles are checked against a registry
    assert lib.cf_destroy(None) == _abi.CF_ERR_HANDLE
    assert lib.cf_destroy(bogus) == _abi.CF_ERR_HANDLE
    p = SimulationParams().to_c()
    assert lib.cf_simulate(bogus, C.byref(p), 1, None) == _abi.CF_ERR_HANDLE
    assert lib.cf_ensemble_simulate(bogus, 1) == _abi.CF_ERR_HANDLE
    assert b"handle" in lib.cf_last_error_string()


def test_cutoff_threshold_equals_sqrt_predicate():
    """s < thr(e)  <=>  sqrtf(s) < e, checked exhaustively over +-3000 floats around thr
    for a spread of radii (the engine's cutoff test, ParticleSimulation.cu:105, :112)."""
    rng = np.random.default_rng(0)
    radii = np.concatenate([rng.uniform(0.005, 0.2, 40), rng.uniform(1, 700, 20),
                            [0.04, 0.05, 0.052, 42.07, 285.0]]).astype(np.float32)
    for e in radii:
        thr = np.float32(cutoff_threshold(float(e)))
        base = thr.view(np.int32)
        s = (base + np.arange(-3000, 3001, dtype=np.int32)).view(np.float32)
        s = s[s >= 0]
        assert np.array_equal(s < thr, np.sqrt(s) < e), e
    assert cutoff_threshold(0.0) == 0.0 and cutoff_threshold(-1.0) == 0.0
    assert cutoff_threshold(float("inf")) == float("inf")


def test_grid_rule():
    p = SimulationParams(radius=0.04, canvasWidth=2.0, canvasHeight=2.0, canvasDepth=2.0,
                         numParticleTypes=8, ratioWithLFO=0.3)
    rbt = np.linspace(-0.9, 0.8, 8).astype(np.float32)
    d = grid_dims(p, rbt, 100000)
    cut = max(np.float32(0.04) * (np.float32(1) + r * np.float32(0.3)) for r in rbt)
    for a in range(3):
        assert 2.0 / d[a] >= cut * (1 + 1 / 1024) - 1e-9
        assert 2.0 / (d[a] + 1) < cut * (1 + 1 / 1024)
    d2 = grid_dims(p, rbt, 100000, cell_divisor=2)
    assert all(2.0 / d2[a] >= cut * (1 + 1 / 1024) / 2 - 1e-9 for a in range(3))
    tiny = grid_dims(p, rbt, 10)  # capped near 8 cells per particle (at least 4096)
    assert tiny[0] * tiny[1] * tiny[2] <= 4096
    with pytest.raises(_abi.CellFlowError):
        grid_dims(p, rbt, 100, cell_divisor=3)
    bad = SimulationParams(canvasWidth=0.0)
    with pytest.raises(_abi.CellFlowError):
        grid_dims(bad, np.zeros(6, np.float32), 100)
