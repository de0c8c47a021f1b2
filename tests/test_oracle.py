"""Pins the CPU oracle (oracle/oracle.cpp) before the engine is compared against it.

The reference ships no tests or golden vectors, so each oracle function is checked against
what the mathematics and the reference's own semantics fix, chosen so that a plausible
mistake (wrong sign, swapped table index, dropped softening, wrong truncation order, a
missed periodic image, an off-by-one in the count ping-pong) fails at least one test:

* closed forms: an isolated particle, two particles, zero table, balance = 1, ratio = 0;
* invariants: pair symmetry of the neighbour relation, permutation equivariance, exact
  periodic-shift equivariance, all-pairs vs grid enumeration;
* brute force in FP64 on pairs away from the cutoff;
* hand-built proximity-graph cases whose answer follows from the reference's selection rule
  (ParticleSimulation.cu:214-246) by inspection;
* the reference's preset files (tests/golden/presets) for the table transform
  (ParticleSimulation.cu:521-531).
"""

import json
import math
import os

import numpy as np
import pytest

from cellflow_amd import configs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

BASE = dict(radius=0.05, delta_t=0.01, friction=0.1, repulsion=64.83, attraction=3.06, k=1.0,
            balance=0.79, force_multiplier=2.33, max_expected_neighbors=400,
            canvas=[1.0, 1.0, 1.0], num_types=2, ratio=0.0)


def phys(**over):
    p = dict(BASE)
    p.update(over)
    return p


def one_step(orc, pos, ptype, table, rbt, ph, vel=None, prev=None, mode=0):
    n = len(ptype)
    vel = np.zeros((n, 3)) if vel is None else vel
    prev = np.zeros(n, np.int32) if prev is None else prev
    return orc.step(np.asarray(pos, np.float64), vel, np.asarray(ptype, np.uint32), prev, ph,
                    np.asarray(table, np.float32), np.asarray(rbt, np.float32), mode=mode)


# ---- closed forms --------------------------------------------------------------------


def test_isolated_particle_friction_only(orc):
    ph = phys()
    v0 = np.array([[0.3, -0.2, 0.1]])
    out = one_step(orc, [[0.5, 0.5, 0.5]], [0], np.ones((2, 2)), [0, 0], ph, vel=v0)
    assert out["count"][0] == 0
    np.testing.assert_allclose(out["vel"][0], v0[0] * 0.1, rtol=1e-6)   # v <- friction * v
    np.testing.assert_allclose(out["pos"][0], 0.5 + v0[0] * 0.1 * 0.01, rtol=1e-6)
    assert np.all(out["acc"][0] == 0.0)


def test_two_particles_direction_and_table_row(orc):
    """Force on i uses table[type_i][type_j] (row = own type) and a positive net value
    pulls i towards j (ParticleSimulation.cu:116, :125-131)."""
    ph = phys()
    table = np.array([[0.0, 0.8], [0.0, 0.0]], np.float32)  # only type 0 feels type 1
    pos = [[0.50, 0.5, 0.5], [0.52, 0.5, 0.5]]
    out = one_step(orc, pos, [0, 1], table, [0, 0], ph)
    assert out["count"].tolist() == [1, 1]                 # both see each other
    assert out["acc"][0][0] > 0 and out["acc"][0][1] == 0  # i pulled toward +x (toward j)
    assert np.all(out["acc"][1] == 0)                      # table[1][0] == 0: j feels nothing
    # magnitude: F = fm * fac * fv * (rep*exp(-k r^2) - att*r), r = sqrt(d^2+1e-4)/R
    d = 0.52 - 0.50
    dist = math.sqrt(d * d + 1e-4)
    r = dist / 0.05
    fac = 1 - (1 - 0.79) * min(0.5 / 400, 1)
    f = 2.33 * fac * 0.8 * (64.83 * math.exp(-r * r) - 3.06 * r) * d / dist
    assert out["acc"][0][0] == pytest.approx(f, rel=1e-6)


def test_softening_and_strict_cutoff(orc):
    """dist = sqrt(s + 1e-4): two particles at separation s with s + 1e-4 = R^2 do not
    interact (strict <), slightly closer they do (ParticleSimulation.cu:105, :112)."""
    R = 0.05
    edge = math.sqrt(R * R - 1e-4)  # separation where sqrt(d^2 + 1e-4) == R
    ph = phys(radius=R)
    for d, expect in [(edge * 0.999, 1), (edge * 1.001, 0)]:
        out = one_step(orc, [[0.3, 0.3, 0.3], [0.3 + d, 0.3, 0.3]], [0, 0], np.ones((2, 2)),
                       [0, 0], ph)
        assert out["count"].tolist() == [expect, expect]


def test_zero_table_is_pure_friction(orc):
    st = configs.make_state(dict(configs.load("c1_plain_10k"), n=1500, canvas=[0.5] * 3))
    ph = configs.physics(dict(configs.load("c1_plain_10k"), canvas=[0.5] * 3))
    vel = np.random.default_rng(1).normal(0, 1, (1500, 3))
    out = one_step(orc, st["pos"], st["type"], np.zeros_like(st["table"]), st["rbt"], ph, vel=vel)
    np.testing.assert_allclose(out["vel"], vel * ph["friction"], rtol=1e-6)
    assert out["count"].sum() > 0  # neighbours still counted


def test_balance_one_ignores_counts(orc):
    st = configs.make_state(dict(configs.load("c1_plain_10k"), n=1500, canvas=[0.5] * 3))
    ph = configs.physics(dict(configs.load("c1_plain_10k"), canvas=[0.5] * 3), balance=1.0)
    a = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    b = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph,
                 prev=np.full(1500, 777, np.int32))
    np.testing.assert_array_equal(a["vel"], b["vel"])
    ph2 = dict(ph, balance=0.5)
    c = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph2)
    d = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph2,
                 prev=np.full(1500, 800, np.int32))
    # density factor saturates at 1: fac = balance = 0.5 exactly for large previous counts
    fac_c = 1 - 0.5 * np.minimum(c["count"] * 0.5 / 400, 1)
    expect = c["acc"] * (0.5 / fac_c)[:, None]
    np.testing.assert_allclose(d["acc"], expect, rtol=1e-12, atol=1e-12)


def test_ratio_zero_makes_radius_uniform(orc):
    ph = phys(ratio=0.0)
    rbt = [0.9, -0.7]
    for tj in (0, 1):
        assert orc.eff_radius(ph, rbt, 0, tj) == np.float32(0.05)
    ph = phys(ratio=0.5)
    assert orc.eff_radius(ph, rbt, 0, 0) == np.float32(np.float32(0.05) * np.float32(1.45))


# ---- invariants ------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c1_small():
    cfg = dict(configs.load("c1_plain_10k"), n=3000, canvas=[0.6, 0.6, 0.6])
    return configs.make_state(cfg), configs.physics(cfg)


def test_pair_symmetry_counts_even(orc, c1_small):
    st, ph = c1_small
    out = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    assert out["count"].sum() % 2 == 0 and out["count"].sum() > 0


def test_grid_matches_all_pairs(orc, c1_small):
    st, ph = c1_small
    a = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph, mode=0)
    b = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph, mode=1)
    np.testing.assert_array_equal(a["count"], b["count"])
    np.testing.assert_allclose(a["vel"], b["vel"], rtol=1e-12, atol=1e-13)


def test_permutation_equivariance(orc, c1_small):
    st, ph = c1_small
    perm = np.random.default_rng(5).permutation(len(st["type"]))
    a = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    b = one_step(orc, st["pos"][perm], st["type"][perm], st["table"], st["rbt"], ph)
    np.testing.assert_array_equal(a["count"][perm], b["count"])
    np.testing.assert_allclose(a["vel"][perm], b["vel"], rtol=1e-12, atol=1e-13)


def test_periodic_shift_equivariance(orc):
    """Shifting every position by L/2 (exact in FP32 for these positions) changes no
    pairwise minimum-image difference, so counts and forces are unchanged."""
    cfg = dict(configs.load("c1_plain_10k"), n=3000)
    st, ph = configs.make_state(cfg), configs.physics(cfg)
    sh = np.mod(st["pos"].astype(np.float64) + 0.5, 1.0)
    assert np.array_equal(sh.astype(np.float32).astype(np.float64), sh)
    a = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    b = one_step(orc, sh, st["type"], st["table"], st["rbt"], ph)
    np.testing.assert_array_equal(a["count"], b["count"])
    np.testing.assert_allclose(a["acc"], b["acc"], rtol=1e-9, atol=1e-9)


def test_counts_match_fp64_brute_force(orc, c1_small):
    """Away from the cutoff (|dist - R| > 1e-6), the FP32 decision equals the FP64 one."""
    st, ph = c1_small
    out = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    x = st["pos"].astype(np.float64)
    L = 0.6
    d = x[None, :, :] - x[:, None, :]
    d -= L * np.round(d / L)
    dist = np.sqrt((d * d).sum(-1) + 1e-4)
    rb = st["rbt"].astype(np.float64)[st["type"]]
    R = 0.05 * (1 + rb * ph["ratio"])
    eff = 0.5 * (R[:, None] + R[None, :])
    np.fill_diagonal(dist, np.inf)
    inside = dist < eff
    safe = np.abs(dist - eff) > 1e-6
    rows_safe = safe.all(axis=1)
    assert rows_safe.mean() > 0.99
    np.testing.assert_array_equal(out["count"][rows_safe], inside.sum(1)[rows_safe])


def test_rollout_carries_counts(orc, c1_small):
    """Two single steps == one two-step rollout (count ping-pong, ParticleSimulation.cu:541-556)."""
    st, ph = c1_small
    n = len(st["type"])
    s1 = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    s2 = orc.step(s1["pos"], s1["vel"], st["type"], s1["count"], ph, st["table"], st["rbt"])
    pos, vel, prev = orc.rollout(st["pos"], np.zeros((n, 3)), st["type"], np.zeros(n, np.int32),
                                 ph, st["table"], st["rbt"], 2)
    np.testing.assert_array_equal(prev, s2["count"])
    np.testing.assert_allclose(pos, s2["pos"], rtol=0, atol=1e-15)


# ---- proximity graph -------------------------------------------------------------------


def test_edges_index_truncation_then_distance(orc):
    """Owner 0 has 20 same-type partners within D; with maxConn 4 the reference keeps the
    first 2*4 = 8 by index (ids 1..8), sorts those by distance and keeps 4.  The partners
    nearest overall (ids 15..20) are never considered."""
    n = 21
    pos = np.zeros((n, 3), np.float32)
    pos[0] = [0.5, 0.5, 0.5]
    for j in range(1, n):
        # ids 1..8 at distances 0.0090 .. 0.0020 (decreasing), ids 9..20 even closer
        r = 0.0100 - 0.0010 * j if j <= 8 else 0.0005 + 0.00001 * j
        ang = 2 * math.pi * j / n
        pos[j] = [0.5 + r * math.cos(ang), 0.5 + r * math.sin(ang), 0.5]
    ptype = np.zeros(n, np.uint32)
    e = orc.edges(pos, ptype, 0.02, 4, mode=0)
    owner0 = e[e[:, 0] == 0, 1].tolist()
    assert owner0 == [8, 7, 6, 5]  # the 4 nearest among ids 1..8, nearest first
    np.testing.assert_array_equal(orc.edges(pos, ptype, 0.02, 4, mode=1), e)


def test_edges_ties_keep_index_order(orc):
    pos = np.array([[0.5, 0.5, 0.5], [0.51, 0.5, 0.5], [0.49, 0.5, 0.5], [0.5, 0.51, 0.5]],
                   np.float32)
    e = orc.edges(pos, np.zeros(4, np.uint32), 0.02, 2, mode=0)
    assert e[e[:, 0] == 0, 1].tolist() == [1, 2] or e[e[:, 0] == 0, 1].tolist() == [2, 1]
    d1 = np.float32(0.51) - np.float32(0.5)
    d2 = np.float32(0.49) - np.float32(0.5)
    if d1 * d1 == d2 * d2:
        assert e[e[:, 0] == 0, 1].tolist() == [1, 2]


def test_edges_same_type_upper_nonperiodic(orc):
    pos = np.array([[0.001, 0.5, 0.5], [0.999, 0.5, 0.5],   # close only across the boundary
                    [0.3, 0.3, 0.3], [0.305, 0.3, 0.3],     # close, same type
                    [0.6, 0.6, 0.6], [0.605, 0.6, 0.6]],    # close, different type
                   np.float32)
    ptype = np.array([0, 0, 1, 1, 2, 3], np.uint32)
    e = orc.edges(pos, ptype, 0.02, 8, mode=0)
    assert e.tolist() == [[2, 3]]  # j > i only; no periodic edge; no cross-type edge


def test_edges_strict_distance(orc):
    d = np.float32(0.02)
    pos = np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]], np.float32)
    s = (pos[1, 0] - pos[0, 0]) ** 2
    e = orc.edges(pos, np.zeros(2, np.uint32), float(d), 4, mode=0)
    assert (len(e) == 1) == bool(s < d * d)


@pytest.mark.parametrize("maxc", [1, 3, 16, 20, 32])
def test_edges_grid_matches_reference_scan(orc, maxc):
    cfg = dict(configs.load("c1_plain_10k"), n=4000, num_types=2, seed=4,
               init=dict(kind="blobs", count=5, sd=0.01))
    st = configs.make_state(cfg)
    for dist in (0.003, 0.02):
        a = orc.edges(st["pos"], st["type"], dist, maxc, mode=0)
        b = orc.edges(st["pos"], st["type"], dist, maxc, mode=1)
        np.testing.assert_array_equal(a, b)
        assert len(a) > 0
        # every kept pair satisfies the predicate, owners ascending, at most maxc each
        assert np.all(a[:, 1] > a[:, 0]) and np.all(np.diff(a[:, 0]) >= 0)
        assert np.bincount(a[:, 0]).max() <= min(maxc, 32)


def test_edges_nonpositive_maxconn(orc):
    pos = np.array([[0.5, 0.5, 0.5], [0.505, 0.5, 0.5]], np.float32)
    assert len(orc.edges(pos, np.zeros(2, np.uint32), 0.02, 0, mode=0)) == 0
    assert len(orc.edges(pos, np.zeros(2, np.uint32), 0.02, -3, mode=0)) == 0


# ---- triangle mesh (ParticleSimulation.cu:279-418, threshold :722-724) -------------------


def test_triangles_square_by_hand(orc):
    """Four same-type particles on a square of side 0.01 (diagonal 0.0141, all within
    1.2 * 0.02) and one far away: owner 0 fans over {1, 2, 3}, owner 1 over {2, 3}."""
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0], [0.01, 0.01, 0], [0.5, 0.5, 0.5]],
                   np.float32)
    for mode in (0, 1):
        t = orc.triangles(pos, np.zeros(5, np.uint32), 0.02, 8, mode=mode)
        assert t.tolist() == [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
    # other types are never neighbours; the hysteresis widens the distance by 1.2
    ty = np.array([0, 1, 0, 0, 0], np.uint32)
    assert orc.triangles(pos, ty, 0.02, 8, mode=0).tolist() == [[0, 2, 3]]
    # reach 0.0095 (below the side): nothing; reach 0.0105 (sides, not diagonals): owner 0
    # has {1, 2}, owners 1 and 2 have one neighbour above them at most
    assert len(orc.triangles(pos, np.zeros(5, np.uint32), 0.0095 / 1.2, 8, mode=0)) == 0
    assert orc.triangles(pos, np.zeros(5, np.uint32), 0.0105 / 1.2, 8, mode=0).tolist() == \
        [[0, 1, 2]]


def test_triangles_lower_neighbours_fill_the_list(orc):
    """Six particles all within reach: owner 3's list holds the first three indices {0, 1, 2}
    (all below it) at maxConn 3, so it owns no triangle though 4 and 5 are near."""
    rng = np.random.default_rng(1)
    pos = (0.5 + 0.002 * rng.random((6, 3))).astype(np.float32)
    t = orc.triangles(pos, np.zeros(6, np.uint32), 0.05, 3, mode=0)
    owners = t[:, 0].tolist()
    assert 3 not in owners and owners.count(0) == 3 and owners.count(1) == 1
    assert t[owners.index(1)].tolist() == [1, 2, 3]


def test_triangles_cap_of_64_in_fan_order(orc):
    """20 particles within reach: owner 0's 19 neighbours give C(19, 2) = 171 pairs, of
    which the first 64 in (a, b) order are kept."""
    rng = np.random.default_rng(2)
    pos = (0.5 + 0.001 * rng.random((20, 3))).astype(np.float32)
    t = orc.triangles(pos, np.zeros(20, np.uint32), 0.05, 32, mode=0)
    own0 = t[t[:, 0] == 0]
    pairs = [(a, b) for a in range(1, 20) for b in range(a + 1, 20)][:64]
    assert [tuple(x) for x in own0[:, 1:].tolist()] == pairs
    counts = np.bincount(t[:, 0], minlength=20)
    assert counts.tolist() == [min(64, (19 - i) * (18 - i) // 2) for i in range(20)]


@pytest.mark.parametrize("maxc", [1, 2, 5, 32, 40])
def test_triangles_grid_matches_reference_scan(orc, maxc):
    cfg = dict(configs.load("c1_plain_10k"), n=3000, num_types=2, seed=5,
               init=dict(kind="blobs", count=5, sd=0.01))
    st = configs.make_state(cfg)
    for dist in (0.003, 0.01):
        a = orc.triangles(st["pos"], st["type"], dist, maxc, mode=0)
        b = orc.triangles(st["pos"], st["type"], dist, maxc, mode=1)
        np.testing.assert_array_equal(a, b)
        if maxc >= 2:
            assert len(a) > 0
        # i < a < b, same type, both within reach, owners ascending, at most 64 each
        assert np.all(a[:, 0] < a[:, 1]) and np.all(a[:, 1] < a[:, 2])
        assert np.all(np.diff(a[:, 0]) >= 0)
        ty = st["type"]
        assert np.all(ty[a[:, 0]] == ty[a[:, 1]]) and np.all(ty[a[:, 0]] == ty[a[:, 2]])
        p = st["pos"].astype(np.float64)
        for k in (1, 2):
            d = np.sqrt(((p[a[:, k]] - p[a[:, 0]]) ** 2).sum(1))
            assert np.all(d < 1.2 * dist * (1 + 1e-5))
        if len(a):
            assert np.bincount(a[:, 0]).max() <= 64


def test_triangle_vertices_normal_and_colour(orc):
    pos = np.array([[0, 0, 0], [0.3, 0, 0], [0, 0.2, 0], [1, 1, 1], [2, 2, 2], [3, 3, 3]],
                   np.float32)
    ty = np.array([4, 0, 0, 1, 1, 1], np.uint32)
    col = np.arange(18, dtype=np.float32).reshape(6, 3) / 10
    v = orc.triangle_vertices(np.array([[0, 1, 2], [3, 4, 5]], np.int32), pos, ty, 3, col)
    assert v.shape == (6, 9)
    np.testing.assert_array_equal(v[:3, :3], pos[:3])
    np.testing.assert_array_equal(v[:3, 3:6], [[0, 0, 1]] * 3)   # (x edge) x (y edge) = +z
    np.testing.assert_array_equal(v[:3, 6:], [col[4 % 3]] * 3)  # the owner's type mod T
    np.testing.assert_array_equal(v[3:, 3:6], 0)                 # collinear: left unnormalised


# ---- binning ---------------------------------------------------------------------------


def test_binning_rule_and_order(orc):
    pos = np.array([[0.0, 0.0, 0.0], [0.999999, 0.5, 0.25], [0.1, 0.9, 0.9], [0.1, 0.9, 0.9]],
                   np.float32)
    cells = orc.bin_cells(pos, [10, 10, 10], [1.0, 1.0, 1.0])
    assert cells.tolist() == [0, 9 + 10 * (5 + 10 * 2), 1 + 10 * (9 + 10 * 9), 1 + 10 * (9 + 10 * 9)]
    order = orc.order(cells)
    assert order.tolist() == [0, 1, 2, 3]
    # clamping: x >= L lands in the last cell, negative in the first
    c2 = orc.bin_cells(np.array([[1.0, -0.1, 0.5]], np.float32), [10, 10, 10], [1.0] * 3)
    assert c2[0] == 9 + 10 * (0 + 10 * 5)


def test_every_interacting_pair_within_stencil(orc):
    """With the engine's grid rule, every pair the predicate accepts lies in adjacent cells
    (periodic): the 27-cell stencil sees all neighbours."""
    from cellflow_amd import SimulationParams, grid_dims
    cfg = dict(configs.load("c2_100k_edges"), n=6000, canvas=[0.8, 0.8, 0.8])
    st, ph = configs.make_state(cfg), configs.physics(cfg)
    dims = grid_dims(SimulationParams.from_physics(ph), st["rbt"], 6000)
    cells = orc.bin_cells(st["pos"], dims, ph["canvas"]).astype(np.int64)
    c = np.stack([cells % dims[0], (cells // dims[0]) % dims[1], cells // (dims[0] * dims[1])], 1)
    x = st["pos"].astype(np.float64)
    d = x[None] - x[:, None]
    d -= 0.8 * np.round(d / 0.8)
    near = np.sqrt((d * d).sum(-1) + 1e-4) < 0.04 * 1.5
    i, j = np.nonzero(near)
    dc = np.abs(c[i] - c[j])
    dc = np.minimum(dc, np.array(dims) - dc)
    ok = np.sqrt((d[i, j] ** 2).sum(-1) + 1e-4) >= max(
        orc.eff_radius(ph, st["rbt"], a, b) for a in range(8) for b in range(8))
    assert np.all((dc <= 1).all(axis=1) | ok)


# ---- force table transform (reference presets) -----------------------------------------


@pytest.mark.parametrize("name", ["settings", "eater", "littlecells", "pulser"])
def test_force_table_transform_on_presets(orc, name):
    d = json.load(open(os.path.join(GOLDEN, "presets", name + ".json")))
    raw = np.array(d["rawForceTable"], np.float32)
    T = d["numParticleTypes"]
    assert raw.size == T * T
    rng_, bias, off = (np.float32(d["forceRange"]), np.float32(d["forceBias"]),
                       np.float32(d["forceOffset"]))
    out = orc.force_table(raw, rng_, bias, off)
    assert np.all(np.abs(out) <= 1.0)
    ref = np.clip(np.tanh(raw.astype(np.float64) * off) * rng_ + bias, -1, 1)
    np.testing.assert_allclose(out, ref, atol=2e-7)
    # the engine's host transform is the same function, bit for bit
    from cellflow_amd import force_table_transform
    np.testing.assert_array_equal(force_table_transform(raw, rng_, bias, off), out)
    # offset 0: every entry is clamp(bias)
    np.testing.assert_array_equal(orc.force_table(raw, rng_, bias, 0.0), np.full_like(out, bias))


# ---- seeded generator (SplitMix64 known answers) ------------------------------------------


def test_splitmix64_known_answers():
    # Steele/Lea/Flood SplitMix64 from state 0: the first three outputs.
    state = np.array([0x9E3779B97F4A7C15, 2 * 0x9E3779B97F4A7C15 % 2**64,
                      3 * 0x9E3779B97F4A7C15 % 2**64], np.uint64)
    out = configs._mix(state)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_ranges():
    cfg = configs.load("c3_1m_clustered")
    st = configs.make_state(dict(cfg, n=20000))
    assert st["pos"].dtype == np.float32 and np.all(st["pos"] >= 0) and np.all(st["pos"] < 6.0)
    assert st["type"].max() == 7 and st["type"].min() == 0
    assert np.all(np.abs(st["rbt"]) < 1)
    ens = configs.make_ensemble(configs.load("c4_ensemble_256"), sims=8)
    for ph, s in ens:
        assert 2000 <= len(s["type"]) <= 50000 and 4 <= ph["num_types"] <= 8
        assert 0.03 <= ph["radius"] <= 0.06


# ---- differentiable-rollout helpers (soft histogram, neighbour lists, frozen rollout) --------


def test_hist_two_particles_known_spline_values(orc):
    """Two particles at distance r: each sees the other once, so h_b = B3(r/w - 1/2 - b).
    The cubic B-spline's values at integer and half-integer offsets are known exactly:
    B3(0) = 2/3, B3(+-1) = 1/6, B3(+-2) = 0, B3(+-1/2) = 23/48, B3(+-3/2) = 1/48.
    (r_max = 1/8 is exact in FP32, the precision the engine's ABI takes it in.)"""
    bins, r_max = 16, 0.125
    w = r_max / bins
    L = [1.0, 1.0, 1.0]
    for k, half in [(5, 0.0), (9, 0.5), (0, 0.0), (14, 0.5)]:
        r = (k + 0.5 + half) * w
        for pos in ([[0.2, 0.3, 0.4], [0.2 + r, 0.3, 0.4]],
                    [[0.99, 0.5, 0.5], [0.99 + r - 1.0, 0.5, 0.5]]):  # across the boundary
            h = orc.hist(np.array(pos), L, bins, r_max)
            want = np.zeros(bins)
            vals = {0: 2 / 3, 1: 1 / 6, -1: 1 / 6} if half == 0.0 else \
                {0: 23 / 48, -1: 23 / 48, 1: 1 / 48, -2: 1 / 48}
            for off, v in vals.items():
                b = k + off if half == 0.0 else k + 1 + off
                if 0 <= b < bins:
                    want[b] = v
            np.testing.assert_allclose(h, want, atol=1e-9, err_msg=f"k={k} half={half}")


def test_hist_partition_of_unity_and_pair_count(orc):
    """sum_b B3(u - b) = 1 whenever all four supporting bins exist (1 <= u <= bins - 2), so
    a pair at such a distance adds exactly 2/N to the histogram's total; a pair beyond
    r_max + 3w/2 (past the last bin's support) adds nothing."""
    bins, r_max = 24, 0.12
    w = r_max / bins
    rng = np.random.default_rng(3)
    for r in rng.uniform(1.5 * w, r_max - 1.5 * w, 20):
        d = rng.normal(size=3)
        d *= r / np.linalg.norm(d)
        pos = np.array([[0.5, 0.5, 0.5], 0.5 + d, [0.5, 0.5, 0.85]])  # third one out of reach
        h = orc.hist(pos, [1.0, 1.0, 1.0], bins, r_max)
        assert abs(h.sum() - 2 / 3) < 1e-9, r
    far = np.array([[0.5, 0.5, 0.5], [0.5, 0.5 + r_max + 1.6 * w, 0.5]])
    assert not orc.hist(far, [1.0, 1.0, 1.0], bins, r_max).any()


def test_hist_lattice_shells(orc):
    """A simple cubic lattice of spacing a in a periodic box: every particle sees 6
    neighbours at a, 12 at a*sqrt(2) and 8 at a*sqrt(3) inside r_max, so the histogram is
    the shell sum, independent of how pairs are enumerated (periodic images included)."""
    m, a = 10, 0.1
    g = (np.arange(m) + 0.25) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    bins, r_max = 20, 0.18
    h = orc.hist(pos, [m * a] * 3, bins, r_max)
    w = r_max / bins
    two = orc.hist(np.array([[0.2, 0.2, 0.2], [0.2 + a, 0.2, 0.2]]), [1.0] * 3, bins, r_max)
    two2 = orc.hist(np.array([[0.2, 0.2, 0.2], [0.2 + a, 0.2 + a, 0.2]]), [1.0] * 3, bins, r_max)
    two3 = orc.hist(np.array([[0.2, 0.2, 0.2], [0.2 + a, 0.2 + a, 0.2 + a]]), [1.0] * 3, bins,
                    r_max)
    assert r_max + 2 * w < 2 * a  # the 2a shell is out of reach
    np.testing.assert_allclose(h, 6 * two + 12 * two2 + 8 * two3, rtol=1e-9, atol=1e-12)


def test_pairs_match_step_counts_and_brute_force(orc, c1_small):
    st, ph = c1_small
    pos32 = st["pos"].astype(np.float32)
    offs, nbr = orc.pairs(pos32, st["type"], ph, st["rbt"])
    out = one_step(orc, st["pos"], st["type"], st["table"], st["rbt"], ph)
    np.testing.assert_array_equal(np.diff(offs), out["count"])
    for i in range(0, len(pos32), 97):  # brute force on a sample of owners
        want = [j for j in range(len(pos32)) if j != i and
                orc.pair_in_range(ph, st["rbt"], pos32[i], pos32[j], int(st["type"][i]),
                                  int(st["type"][j]))]
        assert list(nbr[offs[i]:offs[i + 1]]) == want, i


def _rollout_lists(orc, st, ph, steps):
    pos, vel, prev = st["pos"].astype(np.float64), st["vel"].astype(np.float64), st["prev"]
    lists = []
    for _ in range(steps):
        lists.append(orc.pairs(pos.astype(np.float32), st["type"], ph, st["rbt"]))
        pos, vel, prev = orc.rollout(pos, vel, st["type"], prev, ph, st["table"], st["rbt"], 1,
                                     mode=0)
    return lists, pos


def test_frozen_rollout_equals_rollout_on_its_own_decisions(orc):
    """With the neighbour lists of the rollout's own trajectory, the frozen-decision rollout
    is the ordinary FP64 rollout: same final histogram and loss."""
    cfg = dict(configs.load("c5_adjoint_20k"), n=600, canvas=[0.31, 0.31, 0.31])
    st, ph = configs.make_state(cfg), configs.physics(cfg)
    lists, final = _rollout_lists(orc, st, ph, 6)
    tgt = np.linspace(0.0, 2.0, 16)
    fr = orc.FrozenRollout(st["pos"], st["vel"], st["type"], st["prev"], ph, st["rbt"], lists,
                           16, 0.1, tgt)
    loss, h = fr.loss(st["table"].astype(np.float64), float(np.float32(ph["ratio"])),
                      want_hist=True)
    href = orc.hist(final, ph["canvas"], 16, 0.1)
    np.testing.assert_allclose(h, href, rtol=1e-11, atol=1e-13)
    assert abs(loss - ((href - tgt) ** 2).sum()) < 1e-10 * max(1.0, loss)


def test_frozen_fd_gradient_converges_and_unused_entries_are_zero(orc):
    """Central differences at two step sizes agree (the frozen loss is smooth), and a table
    entry that no interacting pair uses has exactly zero derivative."""
    cfg = dict(configs.load("c5_adjoint_20k"), n=300, num_types=3, canvas=[0.25, 0.25, 0.25])
    st, ph = configs.make_state(cfg), configs.physics(cfg)
    st["type"] = np.where(st["type"] == 2, 1, st["type"]).astype(st["type"].dtype)
    st["type"][0] = 2  # a single particle of type 2: the (2, 2) entry is never used
    lists, _ = _rollout_lists(orc, st, ph, 4)
    tgt = np.linspace(0.0, 2.0, 12)
    fr = orc.FrozenRollout(st["pos"], st["vel"], st["type"], st["prev"], ph, st["rbt"], lists,
                           12, 0.1, tgt)
    tab = st["table"].astype(np.float64)
    g1, r1 = fr.fd_gradient(tab, ph["ratio"], rel_step=1e-4)
    g2, r2 = fr.fd_gradient(tab, ph["ratio"], rel_step=2e-5)
    scale = max(np.abs(g1).max(), abs(r1))
    assert np.abs(g1 - g2).max() < 1e-5 * scale and abs(r1 - r2) < 1e-5 * scale
    uses = np.zeros((3, 3), bool)
    for offs, nbr in lists:
        own = np.repeat(np.arange(len(offs) - 1), np.diff(offs))
        uses[st["type"][own], st["type"][nbr]] = True
    assert not uses.all()
    assert np.all(g1[~uses] == 0.0) and np.all(np.abs(g1[uses]) > 0)


# ---- universe shift (moveUniverse) --------------------------------------------------------


def test_move_dyadic_cases_and_inverse(orc):
    """Dyadic inputs make x + d + L exact, so the wrap's answers are known by hand; shifting
    back by -d restores such positions exactly; results stay in [0, L)."""
    L = [1.0, 2.0, 4.0]
    pos = np.array([[0.25, 1.5, 3.75], [0.75, 0.0, 0.5], [0.0, 1.75, 0.0]], np.float32)
    got = orc.move(pos, 0.5, -0.75, 0.25, L)
    want = np.array([[0.75, 0.75, 0.0], [0.25, 1.25, 0.75], [0.5, 1.0, 0.25]], np.float32)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(orc.move(got, -0.5, 0.75, -0.25, L), pos)
    rng = np.random.default_rng(0)
    x = (rng.random((5000, 3)) * np.array(L)).astype(np.float32)
    y = orc.move(x, 123.4, -0.3, 3.9, L)
    assert np.all(y >= 0) and np.all(y < np.array(L, np.float32))
    # a zero shift of a position already in [0, L) gives fmodf(x + L, L): x up to rounding
    z = orc.move(x, 0.0, 0.0, 0.0, L)
    assert np.all(np.abs(z - x) <= np.spacing(np.float32(2 * 4.0)))


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (700, 2), (3000, 3)])
def test_oracle_under_sanitizers(n, seed):
    """The oracle's step (both enumerations), edges, triangles, pair lists (including a list
    too small for its content) and histogram under AddressSanitizer and
    UndefinedBehaviorSanitizer (a g++ build of oracle/oracle.cpp)."""
    import host_sanitizer
    out = host_sanitizer.run("oracle", n, seed)
    assert out.count("rc=0") >= 8
    assert "pairs (capacity 4) rc=3" in out or n < 100
