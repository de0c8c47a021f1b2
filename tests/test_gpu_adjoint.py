"""GPU: the differentiable rollout (hand-written adjoint) against FP64 finite differences.

The engine's gradient is of the rollout with its discrete decisions (interacting pairs,
neighbour counts) held fixed, evaluated on the rollout's FP64 shadow (the FP32 steps make
the decisions, the continuous state is carried in FP64 along them).  The oracle reproduces
exactly that function: its neighbour lists for step k are computed with the engine's FP32
predicate from the engine's own FP32 state at step k (bit-identical decisions), and its
continuous arithmetic is FP64.  Central finite differences of that FP64 function are the
reference gradient.

Tolerances: every component above 0.1% of the largest within 1e-4 relative of the finite
differences, and every component within 1e-5 of the largest; the loss within 1e-8 relative
and the histogram within 1e-7.  On the full 20k configuration, finite differences with two
step sizes (1e-4 and 3e-5 relative) agree to about 1e-6, which is the floor of this check.

The torch.autograd.Function below wraps the C ABI for the test harness only.
"""

import numpy as np
import pytest
import torch

import oracle
from cellflow_amd import SimulationParams, configs
from helpers import make_sim

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-4     # per component, relative
GRAD_ATOL = 1e-5     # every component, relative to the largest
BIG = 1e-3           # components above this fraction of the largest get the relative check


class RolloutLoss(torch.autograd.Function):
    """loss(table, ratio) of an engine rollout; backward returns the adjoint gradient."""

    @staticmethod
    def forward(ctx, table, ratio, run):
        loss, _, gt, gr = run(table.detach().cpu().numpy(), float(ratio))
        ctx.save_for_backward(torch.from_numpy(gt), torch.tensor(gr, dtype=torch.float64))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gt, gr = ctx.saved_tensors
        return g * gt, g * gr, None


def setup(cfg):
    st = configs.make_state(cfg)
    ph = configs.physics(cfg)
    return st, ph


def target_hist(cfg, steps, bins, r_max):
    st, ph = setup(dict(cfg, seed=cfg.get("target_seed", 1)))
    sim = make_sim(st, ph)
    sim.simulate(SimulationParams.from_physics(ph), steps)
    return sim.pairHistogram(bins, r_max)


def engine_states(st, ph, steps, tuning=None):
    """FP32 positions and previous counts of every pre-step state 0..steps-1."""
    sim = make_sim(st, ph)
    if tuning:
        sim.setTuning(**tuning)
    prm = SimulationParams.from_physics(ph)
    pos = []
    for _ in range(steps):
        pos.append(sim.getParticleData()["pos"].copy())
        sim.simulate(prm, 1)
    return pos


def check_against_fd(cfg, steps, ckpt, bins, r_max, tuning=None, fd_steps=(1e-4,)):
    st, ph = setup(cfg)
    tgt = target_hist(cfg, steps, bins, r_max)
    sim = make_sim(st, ph)
    if tuning:
        sim.setTuning(**tuning)
    loss, h, gt, gr = sim.rolloutGrad(SimulationParams.from_physics(ph), steps, ckpt, bins, r_max,
                                      tgt)
    # oracle: FP64 rollout with the engine trajectory's decisions
    pos = engine_states(st, ph, steps, tuning)
    lists = [oracle.pairs(p, st["type"], ph, st["rbt"]) for p in pos]
    fr = oracle.FrozenRollout(st["pos"], st["vel"], st["type"], st["prev"], ph, st["rbt"], lists,
                              bins, r_max, tgt)
    loss64, h64 = fr.loss(st["table"].astype(np.float64), ph["ratio"], want_hist=True)
    g = np.concatenate([gt.ravel(), [gr]])
    fds = []
    for rel_step in fd_steps:
        fd_t, fd_r = fr.fd_gradient(st["table"].astype(np.float64), ph["ratio"],
                                    rel_step=rel_step, abs_step=1e-2 * rel_step)
        f = np.concatenate([fd_t.ravel(), [fd_r]])
        scale = np.abs(f).max()
        err = np.abs(g - f) / scale
        big = np.abs(f) > BIG * scale
        rel_big = (np.abs(g - f)[big] / np.abs(f)[big]).max()
        print(f"loss engine {loss:.12g} oracle {loss64:.12g}; FD step {rel_step:g}: max grad err "
              f"{err.max():.3g} of max |g|; worst relative of components above {BIG:g} of the "
              f"largest {rel_big:.3g}")
        assert err.max() <= GRAD_ATOL
        assert rel_big <= GRAD_RTOL
        fds.append(f)
    if len(fds) > 1:  # the finite differences' own floor: two step sizes against each other
        a, b = fds[0], fds[-1]
        big = np.abs(a) > BIG * np.abs(a).max()
        fd_floor = (np.abs(a - b)[big] / np.abs(a)[big]).max()
        print(f"finite differences, step {fd_steps[0]:g} against {fd_steps[-1]:g}: worst relative "
              f"{fd_floor:.3g}")
        assert fd_floor <= GRAD_RTOL
    assert abs(loss - loss64) <= 1e-8 * abs(loss64)
    np.testing.assert_allclose(h, h64, rtol=1e-7, atol=1e-12)
    return g


def small_cfg(n=3000, T=4, seed=0):
    cfg = dict(configs.load("c5_adjoint_20k"))
    side = (n / 20000.0) ** (1.0 / 3.0)
    cfg.update(n=n, num_types=T, canvas=[side] * 3, seed=seed)
    return cfg


def test_histogram_matches_oracle():
    cfg = small_cfg()
    st, ph = setup(cfg)
    sim = make_sim(st, ph)
    sim.simulate(SimulationParams.from_physics(ph), 5)
    h = sim.pairHistogram(32, 0.1)
    ref = oracle.hist(sim.getParticleData()["pos"].astype(np.float64), ph["canvas"], 32, 0.1)
    np.testing.assert_allclose(h, ref, rtol=1e-12, atol=1e-14)
    assert abs(h.sum() - ref.sum()) < 1e-9


@pytest.mark.parametrize("steps,ckpt", [(40, 10), (37, 7)])
def test_adjoint_matches_finite_differences_small(steps, ckpt):
    check_against_fd(small_cfg(), steps, ckpt, 16, 0.1)


def test_adjoint_ten_types():
    """The most types a simulation may have: every table-gradient column of the pair kernel's
    per-particle partials is in use (a column left out shows as an error of that column's
    whole size)."""
    check_against_fd(small_cfg(n=2500, T=10), 30, 10, 16, 0.1)


@pytest.mark.parametrize("variant,lanes", [(3, 8), (0, 4)])
def test_adjoint_other_force_kernels(variant, lanes):
    """Variant 0 lists each step's in-range partners in its force kernel during the
    recompute; the other variants leave that to the adjoint's own walk.  Both paths, and
    another lane count, against finite differences."""
    check_against_fd(small_cfg(), 30, 10, 16, 0.1,
                     dict(force_variant=variant, lanes_per_particle=lanes))


def test_adjoint_dense_state_pairs_beyond_the_lds_list():
    """A state dense enough that many particles have more in-range pairs than the pair
    kernel's 64-entry LDS list (they take the windowed path) and others fewer."""
    n = 3000
    cfg = small_cfg(n=n)
    cfg.update(canvas=[0.31] * 3)
    st, ph = setup(cfg)
    offs, _ = oracle.pairs(st["pos"], st["type"], ph, st["rbt"])
    per = np.diff(offs)
    assert (per > 64).sum() > n // 10 and (per <= 64).sum() > n // 10, (per.min(), per.max())
    g = check_against_fd(cfg, 12, 5, 16, 0.1)
    # and the windowed path is as deterministic as the listed one
    tgt = target_hist(cfg, 12, 16, 0.1)
    _, _, gt, gr = make_sim(st, ph).rolloutGrad(SimulationParams.from_physics(ph), 12, 5, 16, 0.1,
                                                tgt)
    assert np.concatenate([gt.ravel(), [gr]]).tobytes() == g.tobytes()


def test_checkpoint_interval_and_repeat_bit_identical():
    """The checkpoint interval changes nothing, repeats are identical.  Intervals 1 and 2
    reuse the per-step pair lists soonest (the shadow steps run on their own stream while
    the next FP32 step lists into a buffer), and 7 leaves a short last segment."""
    cfg = small_cfg(n=2000)
    st, ph = setup(cfg)
    tgt = np.linspace(0.5, 3.0, 16)
    prm = SimulationParams.from_physics(ph)
    outs = []
    for ckpt in (5, 30, 1, 2, 7, 5):
        sim = make_sim(st, ph)
        loss, h, gt, gr = sim.rolloutGrad(prm, 30, ckpt, 16, 0.1, tgt)
        outs.append((loss, gt.tobytes(), gr))
    assert all(o == outs[0] for o in outs), [o[0] for o in outs]


def test_autograd_function_wraps_adjoint():
    cfg = small_cfg(n=2000)
    st, ph = setup(cfg)
    tgt = np.linspace(0.5, 3.0, 16)

    def run(table, ratio):
        s2 = dict(st, table=table.astype(np.float32))
        sim = make_sim(s2, ph)
        return sim.rolloutGrad(SimulationParams.from_physics(dict(ph, ratio=ratio)), 20, 10, 16,
                               0.1, tgt)

    table = torch.tensor(st["table"], dtype=torch.float64, requires_grad=True)
    ratio = torch.tensor(ph["ratio"], dtype=torch.float64, requires_grad=True)
    loss = RolloutLoss.apply(table, ratio, run)
    gt, gr = torch.autograd.grad(3.0 * loss, [table, ratio])
    _, _, gt0, gr0 = run(st["table"], ph["ratio"])
    np.testing.assert_array_equal(gt.numpy(), 3.0 * gt0)
    assert gr.item() == 3.0 * gr0


def test_adjoint_full_config_against_finite_differences():
    """The 20,000-particle, 200-step configuration with checkpoints every 20 steps, against
    central differences at two step sizes (which must agree with each other too)."""
    cfg = configs.load("c5_adjoint_20k")
    check_against_fd(cfg, cfg["steps"], cfg["checkpoint_every"], cfg["bins"], cfg["r_max"],
                     fd_steps=(1e-4, 3e-5))


def test_shadow_drift_is_reported():
    """A rollout whose FP64 shadow strays from the FP32 trajectory beyond the histogram's
    cell slack (here: a chaotic 10k-workload rollout at dt 0.01 for 300 steps) is refused
    with CF_ERR_STATE rather than answered with a wrong loss."""
    from cellflow_amd import CellFlowError, _abi
    cfg = configs.load("c1_plain_10k")
    st, ph = setup(cfg)
    sim = make_sim(st, ph)
    with pytest.raises(CellFlowError) as e:
        sim.rolloutGrad(SimulationParams.from_physics(ph), 300, 50, 16, 0.1, np.ones(16))
    assert e.value.code == _abi.CF_ERR_STATE and "shadow" in str(e.value)
