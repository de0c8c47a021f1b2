"""Distribution of the seeded initial state (replacement for initCurandKernel /
initializeParticlesKernel, ParticleSimulation.cu:21-66).  No bit parity with curand is
possible, so the generator the engine's k_init follows bit for bit (checked on the GPU in
tests/test_gpu_step.py and below) is held to the properties the reference's draw has:
uniform positions in the central spawn cube (:45-55) and uniform types (:62)."""

import numpy as np
import pytest

from cellflow_amd import ParticleSimulation, configs

N = 1 << 20


def chi2(counts, expected):
    return float(((counts - expected) ** 2 / expected).sum())


# chi-square upper quantiles at p = 1e-6 by degrees of freedom (scipy.stats.chi2.isf), so a
# correct generator fails one of these about once in 1e6 runs
CHI2_Q = {2: 27.63, 5: 35.89, 9: 44.81, 15: 56.49}


def test_uniform_moments_and_bounds():
    u = configs.uniform_f32(5, configs.S_POS, N).astype(np.float64)
    assert u.min() >= 0.0 and u.max() < 1.0
    # 24-bit grid: every value is k / 2^24
    assert np.all(u * 2.0 ** 24 == np.floor(u * 2.0 ** 24))
    sd = np.sqrt(1.0 / 12.0 / N)
    assert abs(u.mean() - 0.5) < 5 * sd
    assert abs(u.var() - 1.0 / 12.0) < 5 * np.sqrt(1.0 / 180.0 / N)
    hist = np.bincount((u * 16).astype(np.int64), minlength=16)
    assert chi2(hist, N / 16) < CHI2_Q[15]


def test_streams_and_seeds_independent():
    a = configs.uniform_f32(5, configs.S_POS, N).astype(np.float64)
    b = configs.uniform_f32(5, configs.S_TYPE, N).astype(np.float64)
    c = configs.uniform_f32(6, configs.S_POS, N).astype(np.float64)
    for x, y in ((a, b), (a, c), (a[:-1], a[1:])):  # also successive draws of one stream
        assert abs(np.corrcoef(x, y)[0, 1]) < 5 / np.sqrt(N)
    # an offset continues the same stream
    np.testing.assert_array_equal(configs.uniform_f32(5, configs.S_POS, 100, offset=1000),
                                  configs.uniform_f32(5, configs.S_POS, 1100)[1000:])


@pytest.mark.parametrize("T", [1, 3, 10])
def test_type_draw_uniform(T):
    t = np.minimum((configs.uniform_f32(9, configs.S_TYPE, N) * np.float32(T)).astype(np.uint32),
                   T - 1)
    hist = np.bincount(t, minlength=T)
    assert len(hist) == T
    if T > 1:
        assert chi2(hist, N / T) < CHI2_Q[T - 1]


@pytest.mark.gpu
def test_engine_init_spawn_region_and_moments():
    """Spawn region min(spawn, canvas) per axis, centred (:45-55): with canvas (1, 2, 3) and
    spawn 1.5 the x range is the whole axis and y, z are central bands of width 1.5."""
    n, T = N, 6
    sim = ParticleSimulation(n, T)
    canvas = np.array([1.0, 2.0, 3.0], np.float32)
    sim.initializeParticles(21, tuple(canvas.tolist()), 1.5)
    p = sim.getParticleData()
    w = np.minimum(np.float32(1.5), canvas)
    lo = (canvas - w) * np.float32(0.5)
    u = configs.uniform_f32(21, configs.S_POS, 3 * n).reshape(n, 3)
    np.testing.assert_array_equal(p["pos"], lo + u * w)  # the generator, bit for bit
    pos = p["pos"].astype(np.float64)
    assert np.all(pos >= lo) and np.all(pos <= lo + w)
    sd = w.astype(np.float64) * np.sqrt(1.0 / 12.0 / n)
    assert np.all(np.abs(pos.mean(axis=0) - (lo + w / 2)) < 5 * sd)
    hist = np.bincount(p["ptype"], minlength=T)
    assert len(hist) == T and chi2(hist, n / T) < CHI2_Q[T - 1]
    assert not p["vel"].any() and not np.asarray(sim.getNeighborCounts()).any()
