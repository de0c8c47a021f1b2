"""The allocation-free pair histogram (cf_pair_histogram_noalloc) against cf_pair_histogram,
both through the C ABI: bit for bit on two calls in a row (the second with fewer particles
and bins, so it reuses what the first left in the device scratch array), and its size and
state errors (raised on the host, before anything runs)."""

import ctypes as C

import numpy as np
import pytest

from cellflow_amd import CellFlowError, ParticleSimulation, _abi
from helpers import make_sim, small_cfg, state_of

pytestmark = pytest.mark.gpu


def hist_noalloc(sim, bins, r_max):
    h = np.zeros(bins)
    _abi.check(_abi.lib().cf_pair_histogram_noalloc(sim._h, int(bins), float(r_max),
                                                    h.ctypes.data_as(C.POINTER(C.c_double))))
    return h


def hist_alloc(sim, bins, r_max):
    h = np.zeros(bins)
    _abi.check(_abi.lib().cf_pair_histogram(sim._h, int(bins), float(r_max),
                                            h.ctypes.data_as(C.POINTER(C.c_double))))
    return h


def test_noalloc_histogram_equals_the_allocating_one():
    cases = [(small_cfg(n=20000, side=1.0), 32, 0.1),   # the 20k rollout workload's size
             (small_cfg(n=3000, side=0.5, T=3), 16, 0.08)]
    for cfg, bins, r_max in cases:
        st, phys, params = state_of(cfg)
        sim = make_sim(st, phys)
        sim.simulate(params, 3)
        got = hist_noalloc(sim, bins, r_max)
        ref = hist_alloc(sim, bins, r_max)
        assert ref.sum() > 0
        assert got.tobytes() == ref.tobytes()
        sim.close()


def test_noalloc_histogram_errors():
    st, phys, params = state_of(small_cfg(n=2000, side=0.4))
    sim = make_sim(st, phys)
    with pytest.raises(CellFlowError) as e:  # no parameters yet: nothing is binned
        hist_noalloc(sim, 16, 0.1)
    assert e.value.code == _abi.CF_ERR_STATE
    sim.simulate(params, 0)  # bins the state without stepping
    assert hist_noalloc(sim, 16, 0.1).tobytes() == hist_alloc(sim, 16, 0.1).tobytes()
    moved = sim.getParticleData()
    moved["pos"] = (moved["pos"] + np.float32(0.013)) % np.float32(0.4)
    sim.setParticles(moved)  # a load re-bins the new state
    assert hist_noalloc(sim, 16, 0.1).tobytes() == hist_alloc(sim, 16, 0.1).tobytes()
    sim.close()
    big = ParticleSimulation(200_000, 2)  # 64 bins x 200k doubles exceed the 32 MiB array
    with pytest.raises(CellFlowError) as e:
        hist_noalloc(big, 64, 0.1)
    assert e.value.code == _abi.CF_ERR_ARG and "scratch" in str(e.value)
    big.close()
