"""The partition pre-pass oracle (oracle/pd0.c) pinned against the values the
reference captured from the instrumented C encoder on its gradient-64
identity content (rust/crates/svtav1-encoder/src/pd0.rs:2229-2780): the
variance map, qp threshold scaling, lambda, partition / skip rates, the
LVL_0 / LVL_1 / LVL_5 (with subres) / LVL_6 per-block costs, the variance cap,
the detector and the decided trees; plus closed forms (variance of a
uniform / two-level block, the LVL_6 formula) and brute force of the tree
pick on small cost sets."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle'))
import oracle as O  # noqa: E402


def gradient64():
    """The identity-harness gradient (pd0.rs:2231): ((r*255)/h) ^ ((c*3) & 0x3f)."""
    r = np.arange(64)[:, None]
    c = np.arange(64)[None, :]
    return (((r * 255) // 64) ^ ((c * 3) & 0x3f)).astype(np.uint8)


# pd0.rs:2242 C_GRADIENT64_VARS (instrumented library, MDBG sb_var)
C_VARS = [5425, 1343, 1353, 1733, 1893, 336, 341, 340, 338, 645, 773, 837, 901, 645, 773, 837, 901,
          645, 773, 837, 901, 79, 163, 395, 83, 79, 487, 155, 83, 197, 503, 181, 325, 357, 171, 469,
          229, 197, 1099, 1717, 325, 1573, 1047, 661, 1957, 197, 503, 181, 325, 357, 171, 469, 229,
          197, 1099, 1717, 325, 1573, 1047, 661, 1957, 197, 503, 181, 325, 357, 171, 469, 229, 197,
          1099, 1717, 325, 1573, 1047, 661, 1957, 197, 503, 181, 325, 357, 171, 469, 229]


def test_variance_map_matches_c_capture():
    assert O.b64_variance(gradient64(), 0, 0).tolist() == C_VARS


def test_variance_map_closed_forms():
    # uniform: zero everywhere; an SB whose even rows alternate two levels
    # column-wise: every 8x8 has mean (a+b)/2 and variance ((a-b)/2)^2
    assert not O.b64_variance(np.full((64, 64), 77, np.uint8), 0, 0).any()
    p = np.zeros((64, 64), np.uint8)
    p[:, 0::2], p[:, 1::2] = 10, 30
    v = O.b64_variance(p, 0, 0)
    assert (v == 100).all()
    # odd rows are never read (even-row subsampling)
    q = p.copy()
    q[1::2, :] = 255
    assert np.array_equal(O.b64_variance(q, 0, 0), v)
    # the map of an SB at an offset reads only that SB
    big = np.zeros((128, 192), np.uint8)
    big[64:, 128:] = gradient64()
    assert O.b64_variance(big, 128, 64).tolist() == C_VARS


def test_scaling_lambda_rates_match_c():
    assert O.qp_th_scaling(20) == (20, 63)
    assert O.qp_th_scaling(40) == (40, 63)
    assert O.qp_th_scaling(55) == (9146, 10000)
    L = O.pd0lib()
    assert [L.or_kf_full_lambda_8bit(q, p) for q, p in ((80, 20), (160, 40), (220, 55))] == [25650, 248207, 1527856]
    split, none, skip0 = O.pd0_default_rates()
    assert (skip0, none) == (26, 400)
    assert [2 * s for s in split[1:]] == [4040, 2930, 2390]   # doubled at 16 / 32 / 64
    assert split[1:] == [2020, 1465, 1195]                     # LVL_1 (undoubled) split rates


def test_cap_and_detector_match_c():
    L = O.pd0lib()
    v = O.b64_variance(gradient64(), 0, 0)
    assert [L.or_max_block_size_allintra(int(v[0]), q) for q in (20, 40, 55)] == [32, 32, 64]
    assert [L.or_pd0_detector_demotes(v, q) for q in (20, 40, 55)] == [0, 1, 1]
    u = O.b64_variance(np.full((64, 64), 128, np.uint8), 0, 0)
    assert L.or_pd0_detector_demotes(u, 40) == 1
    assert L.or_max_block_size_allintra(0, 20) == 64


def test_lvl6_costs_match_c_and_formula():
    L = O.pd0lib()
    v = O.b64_variance(gradient64(), 0, 0)
    for sq, ox, oy, cost in [(32, 0, 0, 1382), (16, 0, 0, 294), (8, 0, 0, 87), (8, 8, 0, 89), (8, 0, 8, 89),
                             (8, 8, 8, 89), (16, 16, 0, 294), (16, 0, 16, 313), (16, 16, 16, 320), (32, 32, 0, 1382)]:
        assert L.or_lvl6_cost(v, sq, ox, oy, 20) == cost, (sq, ox, oy)
    # zero variance: the unbiased area (bias 1000)
    z = np.zeros(85, np.uint16)
    assert [L.or_lvl6_cost(z, s, 0, 0, 40) for s in (64, 32, 16, 8)] == [4096, 1024, 256, 64]


def test_lvl0_block_costs_match_c():
    y = gradient64()
    cfg = O.pd0_cfg(0, 20, 80)
    assert cfg.lambda_ == 25650
    for sq, ox, oy, cost in [(32, 0, 0, 26185862), (16, 0, 0, 8396609), (8, 0, 0, 2143413), (8, 8, 0, 1990844),
                             (8, 0, 8, 2225589), (8, 8, 8, 2168757), (16, 16, 0, 6559425), (16, 0, 16, 8443329),
                             (16, 16, 16, 8792001), (32, 32, 0, 28871046), (32, 0, 32, 22111622),
                             (32, 32, 32, 22521222)]:
        assert O.pd0_block_cost(y, cfg, 0, sq, ox, oy, subres=0)[0] == cost, (sq, ox, oy)


def test_lvl5_block_costs_match_c_q40():
    y = gradient64()
    cfg = O.pd0_cfg(5, 40, 160)
    for sq, ox, oy, cost in [(32, 0, 0, 187677438), (16, 0, 0, 48981821), (8, 0, 0, 9695714), (8, 8, 0, 11371661),
                             (8, 0, 8, 16542374), (8, 8, 8, 20538852), (16, 16, 0, 41852989),
                             (32, 32, 0, 190877950), (32, 0, 32, 181407102), (32, 32, 32, 183892222),
                             (16, 48, 16, 53455823)]:
        assert O.pd0_block_cost(y, cfg, 5, sq, ox, oy)[0] == cost, (sq, ox, oy)


def test_lvl5_block_costs_with_subres_match_c_q55():
    y = gradient64()
    cfg = O.pd0_cfg(5, 55, 220)
    c64, state = O.pd0_block_cost(y, cfg, 5, 64, 0, 0)
    assert c64 == 1708208432 and state == 1   # the 64x64 DC prediction passes the odd/even check
    for sq, ox, oy, cost in [(32, 0, 0, 522128378), (16, 0, 0, 137213980), (16, 16, 0, 135635996),
                             (16, 0, 16, 232128024), (16, 16, 16, 194500372), (32, 32, 0, 594523898),
                             (32, 0, 32, 475114621), (32, 32, 32, 469165693)]:
        assert O.pd0_block_cost(y, cfg, 5, sq, ox, oy, subres=state)[0] == cost, (sq, ox, oy)


@pytest.mark.parametrize('qp,qidx,expect', [
    (55, 220, [(64, 0, 0, 1791569177), (32, 0, 0, 526486441), (32, 32, 0, 572301943), (16, 0, 0, 146206469),
               (8, 0, 0, 44014180), (8, 8, 0, 35188942), (8, 0, 8, 37535984), (8, 8, 8, 60514499)]),
    (40, 160, [(64, 0, 0, 1176293547), (32, 0, 0, 230378290), (16, 0, 0, 62496975), (8, 0, 0, 16077204)]),
    (20, 80, [(64, 0, 0, 903280295), (32, 0, 0, 51245980), (16, 0, 0, 14528276), (8, 0, 0, 3483565),
              (8, 8, 0, 3484388)]),
])
def test_lvl1_block_costs_match_c(qp, qidx, expect):
    y = gradient64()
    cfg = O.pd0_cfg(1, qp, qidx, coeff_rate_est_lvl=1)
    for sq, ox, oy, cost in expect:
        assert O.pd0_block_cost(y, cfg, 1, sq, ox, oy)[0] == cost, (qp, sq, ox, oy)


def test_trees_match_c():
    y, u = gradient64(), np.full((64, 64), 128, np.uint8)
    # eff-M9 (LVL_6 / LVL_5 by the detector): pd0.rs:2700
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(5, 20, 80))[1]) == [16] * 16
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(5, 40, 160))[1]) == [32] * 4
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(5, 55, 220))[1]) == [64]
    assert O.leaf_sizes(O.pd0_sb(u, O.pd0_cfg(5, 40, 160))[1]) == [64]
    # LVL_0 (10-bit forced full-RD level): pd0.rs:2347
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(0, 20, 80))[1]) == [32] * 4
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(0, 55, 220))[1]) == [64]
    # LVL_1 (presets 2..8): pd0.rs:2770
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(1, 20, 80))[1]) == [32] * 4
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(1, 40, 160))[1]) == [32] * 4
    assert O.leaf_sizes(O.pd0_sb(y, O.pd0_cfg(1, 55, 220))[1]) == [64]
    assert O.leaf_sizes(O.pd0_sb(u, O.pd0_cfg(1, 40, 160))[1]) == [64]


def test_node_records_follow_the_walk():
    """The records the depth refinement reads: the 64x64 untested when the
    variance cap removes it, its four 32x32 tested with the captured costs,
    children of a kept parent visited up to the early exit."""
    y = gradient64()
    nodes, lm, mode = O.pd0_sb(y, O.pd0_cfg(0, 20, 80))
    assert mode == 0
    f = nodes['flags']
    assert f[0] & O.PD0_VISITED and not f[0] & O.PD0_TESTED and f[0] & O.PD0_SPLIT
    assert nodes['cost'][1:5].tolist() == [26185862, 28871046, 22111622, 22521222]
    assert all(f[1:5] & O.PD0_TESTED) and not any(f[1:5] & O.PD0_SPLIT)
    # 16x16 (0,0) was costed under the first 32x32 (captured 8396609)
    assert f[5] & O.PD0_TESTED and nodes['cost'][5] == 8396609


def _pick_brute(costs, split_bits, lam, mode, max_sq, min_sq):
    """Independent statement of the parent-first walk over a precomputed
    cost table {(sq, x, y): cost}: returns the leaf list in coding order."""
    def rd(r):
        return (r * lam + 256) >> 9

    def go(sq, x, y):
        tested = min_sq <= sq <= max_sq
        pc = costs[(sq, x, y)] if tested else None
        if sq <= min_sq:
            return pc, [sq]
        sb = {8: 0, 16: 1, 32: 2, 64: 3}[sq]
        sc = 0 if mode == 6 else rd(split_bits[sb] * (1 if mode == 1 else 2))
        leaves = []
        h = sq // 2
        for i in range(4):
            if mode != 6 and pc is not None and pc * (50 if i == 0 else 1000) <= sc * 1000:
                return pc, [sq]
            c, l = go(h, x + (i & 1) * h, y + (i >> 1) * h)
            sc += c
            leaves += l
        if pc is not None and pc <= sc:
            return pc, [sq]
        return sc, leaves

    return go(64, 0, 0)[1]


@pytest.mark.parametrize('seed', range(6))
def test_tree_pick_equals_brute_force_on_random_content(seed):
    rng = np.random.default_rng(seed)
    # smooth blocks with a few busy quadrants, so every depth wins somewhere
    y = np.clip(128 + np.cumsum(rng.integers(-3, 4, (64, 64)), axis=1), 0, 255).astype(np.uint8)
    qx, qy = rng.integers(0, 4, 2) * 16
    y[qy:qy + 16, qx:qx + 16] = rng.integers(0, 256, (16, 16))
    for level, qp, qidx in ((0, 32, 128), (1, 40, 160), (5, 28, 112)):
        cfg = O.pd0_cfg(level, qp, qidx)
        nodes, lm, mode = O.pd0_sb(y, cfg)
        costs = {}
        state = 0 if level == 0 else 255
        max_sq = 64 if level == 1 else O.pd0lib().or_max_block_size_allintra(int(O.b64_variance(y, 0, 0)[0]), qp)
        if mode == 5 and max_sq == 64:
            costs[(64, 0, 0)], state = O.pd0_block_cost(y, cfg, 5, 64, 0, 0)
        for sq in (64, 32, 16, 8):
            for yy in range(0, 64, sq):
                for xx in range(0, 64, sq):
                    if (sq, xx, yy) not in costs:
                        costs[(sq, xx, yy)] = O.pd0_block_cost(y, cfg, mode, sq, xx, yy, subres=state)[0]
        split, _, _ = O.pd0_default_rates()
        assert O.leaf_sizes(lm) == _pick_brute(costs, split, cfg.lambda_, mode, max_sq, 8), (level, seed)
