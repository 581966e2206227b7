"""Pin the oracle's deblocking and CDEF: the reference's known answers
(dsp/cdef.rs:1222-1300, golden_parity.rs:1064-1140), the spec direction
table, flat-region identities and search-level invariants (CPU)."""
import ctypes

import numpy as np
import pytest


def test_cdef_constrain_and_adjust_known_values(oracle):
    L = oracle.lflib()
    assert L.or_cdef_constrain(1000, 0, 6) == 0
    assert L.or_cdef_constrain(5, 4, 4) == 3 and L.or_cdef_constrain(-5, 4, 4) == -3
    assert L.or_cdef_constrain(0x7f7f - 128, 15, 6) == 0
    assert L.or_cdef_constrain(0x4000 - 128, 15, 6) == 0
    assert L.or_cdef_constrain(2, 15, 3) == 2
    assert L.or_cdef_adjust_strength(12, 0) == 0
    assert L.or_cdef_adjust_strength(12, 63) == 3
    assert L.or_cdef_adjust_strength(12, 1 << 18) == 12


def test_cdef_find_dir_spec_semantics(oracle):
    c = np.array([[x * 30 for x in range(8)] for _ in range(8)], np.uint16)
    assert oracle.cdef_find_dir(c)[0] == 6 and oracle.cdef_find_dir(c)[1] > 0
    assert oracle.cdef_find_dir(c.T.copy())[0] == 2
    assert oracle.cdef_find_dir(np.full((8, 8), 77, np.uint16)) == (oracle.cdef_find_dir(np.full((8, 8), 77, np.uint16))[0], 0)
    # the eight AV1 directions: lines of equal samples along each direction's step
    steps = [(-1, 1), (-0.5, 1), (0, 1), (0.5, 1), (1, 1), (1, 0.5), (1, 0), (1, -0.5)]
    for d, (dy, dx) in enumerate(steps):
        img = np.zeros((8, 8))
        for i in range(8):
            for j in range(8):
                # value constant along (dy, dx): project onto the normal
                img[i, j] = 128 + 40 * np.sin(0.9 * (i * dx - j * dy))
        got, var = oracle.cdef_find_dir(np.round(img).astype(np.uint16))
        assert got == d, (d, got)


def test_cdef_direction_table_spec(oracle):
    """Decode the filter's tap offsets through an impulse: a single non-zero
    neighbour at each spec (dy, dx) offset moves only the taps of the
    directions that use it (spec 7.15.3 Cdef_Directions)."""
    spec = [[(-1, 1), (-2, 2)], [(0, 1), (-1, 2)], [(0, 1), (0, 2)], [(0, 1), (1, 2)],
            [(1, 1), (2, 2)], [(1, 0), (2, 1)], [(1, 0), (2, 0)], [(1, 0), (2, -1)]]
    L = oracle.lflib()
    S = 144
    for d in range(8):
        for k in range(2):
            dy, dx = spec[d][k]
            buf = np.full(144 * 20, 100, np.uint16)
            c0 = 8 * S + 16
            buf[c0 + dy * S + dx] = 140
            out = np.zeros(1, np.int32)
            # 1x1 "block" (bw = bh = 1): primary only, strength 15, large damping
            L.or_cdef_filter_block(out.ctypes.data, 1, buf[c0:].ctypes.data, 15, 0, d, 6, 6, 1, 1, 0, 1)
            assert out[0] > 100, (d, k, out[0])


def test_cdef_flat_identity(oracle):
    L = oracle.lflib()
    S = 144
    buf = np.full(S * 20, 0x7f7f, np.uint16)
    base = 3 * S + 8
    for r in range(8):
        buf[base + r * S: base + r * S + 8] = 77
    out = np.zeros(64, np.int32)
    L.or_cdef_filter_block(out.ctypes.data, 8, buf[base:].ctypes.data, 15, 4, 3, 6, 6, 8, 8, 0, 1)
    assert np.all(out == 77)
    L.or_cdef_filter_block(out.ctypes.data, 8, buf[base:].ctypes.data, 0, 0, 0, 3, 3, 8, 8, 0, 1)
    assert np.all(out == 77)


def test_lf_thresholds_closed_form(oracle):
    L = oracle.lflib()
    for level in range(64):
        for sh in range(8):
            a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            L.or_lf_thresholds(level, sh, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
            bil = level >> ((sh > 0) + (sh > 4))
            if sh > 0:
                bil = min(bil, 9 - sh)
            bil = max(bil, 1)
            assert (a.value, b.value, c.value) == (2 * (level + 2) + bil, bil, level >> 4)


def test_deblock_flat_unchanged_and_step_smoothed(oracle):
    L = oracle.lflib()
    for bd in (8, 10):
        dt = np.uint8 if bd == 8 else np.uint16
        p = np.full((16, 16), 128 << (bd - 8), dt)
        o = p.copy()
        for level in (1, 8, 32, 63):
            for vert, off in ((1, 2 * 16 + 8), (0, 8 * 16 + 3)):
                for ln in (4, 6, 8, 14):
                    L.or_lpf(p.ctypes.data, int(bd > 8), bd, off, 16, vert, ln, level, 0)
        assert np.array_equal(p, o)
        # a small step across a vertical edge is smoothed by every length
        for ln in (4, 6, 8, 14):
            q = np.zeros((16, 16), dt)
            q[:, 8:] = 6 << (bd - 8)
            q[:, :8] = 2 << (bd - 8)
            L.or_lpf(q.ctypes.data, int(bd > 8), bd, 2 * 16 + 8, 16, 1, ln, 40, 0)
            assert q[2, 7] > (2 << (bd - 8)) and q[2, 8] < (6 << (bd - 8)), ln
            assert np.array_equal(q[6:], np.where(np.arange(16) >= 8, 6, 2)[None, :].repeat(10, 0) << (bd - 8))


def test_cdef_search_skip_list_partitions_the_sse(oracle):
    """The search list drops all-skip 8x8s (enc/deblock.rs:652): at 8 bits
    (no per-plane shift) the rows of complementary skip maps sum to the
    unskipped rows, and an all-skip map gives zero rows."""
    rng = np.random.default_rng(5)
    H, W = 72, 136
    ry = rng.integers(0, 256, (H, W)).astype(np.uint8)
    sy = np.clip(ry.astype(int) + rng.integers(-9, 10, ry.shape), 0, 255).astype(np.uint8)
    ru, rv = (rng.integers(0, 256, (H // 2, W // 2)).astype(np.uint8) for _ in range(2))
    su, sv = (np.clip(p.astype(int) + rng.integers(-5, 6, p.shape), 0, 255).astype(np.uint8) for p in (ru, rv))
    blk = rng.uniform(size=(H // 8, W // 8)) < 0.5
    skip_a = np.repeat(np.repeat(blk, 2, 0), 2, 1).astype(np.uint8)
    skip_b = (1 - skip_a).astype(np.uint8)
    fs = np.arange(64)
    for fbr in range(2):
        for fbc in range(3):
            args = ((ry, ru, rv), (sy, su, sv), W, H, 8, 120, 1, fbr, fbc, fs)
            full = oracle.cdef_search_fb(*args)
            a = oracle.cdef_search_fb(*args, skip=skip_a)
            b = oracle.cdef_search_fb(*args, skip=skip_b)
            none = oracle.cdef_search_fb(*args, skip=np.ones_like(skip_a))
            assert np.array_equal(a[0] + b[0], full[0]) and np.array_equal(a[1] + b[1], full[1])
            assert not none[0].any() and not none[1].any()
            assert a[0].any() or b[0].any()


@pytest.mark.parametrize('bd', [8, 10])
def test_deblock_lines_mirror_symmetric(oracle, bd):
    """Every AV1 deblocking filter treats the p and q sides alike: filtering a
    reversed line gives the reversed result.  Exact for the 6/8/14-tap
    smoothing paths and filter4's inner taps (its rounding pair (f+4)>>3 /
    (f+3)>>3 is antisymmetric away from the clamps); filter4's outer-tap
    step ROUND_POWER_OF_TWO(f1, 1) rounds halves up, so p1 / q1 may differ
    by one.  A swapped tap weight, side or index breaks this."""
    L = oracle.lflib()
    rng = np.random.default_rng(9 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    sh = bd - 8
    hits = 0
    for ln in (4, 6, 8, 14):
        for trial in range(400):
            base = int(rng.integers(40, 200)) << sh
            step = int(rng.integers(-12, 13)) << sh
            noise = rng.integers(-2, 3, ln) << sh if trial % 2 else np.zeros(ln, int)
            line = np.array([base + (step if k >= ln // 2 else 0) for k in range(ln)]) + noise
            buf = np.zeros((4, 16), dt)
            buf[:, 8 - ln // 2:8 + ln // 2] = line
            rev = buf[:, ::-1].copy()  # edge now between columns 7 and 8 of the mirrored row
            level = int(rng.integers(8, 64))
            a, b = buf.copy(), rev.copy()
            L.or_lpf(a.ctypes.data, int(bd > 8), bd, 8, 16, 1, ln, level, 0)
            L.or_lpf(b.ctypes.data, int(bd > 8), bd, 8, 16, 1, ln, level, 0)
            diff = np.abs(a[:, ::-1].astype(int) - b.astype(int))
            diff[:, [6, 9]] = np.maximum(diff[:, [6, 9]] - 1, 0)  # filter4 outer taps: +-1
            assert not diff.any(), (ln, line.tolist(), level)
            hits += int(not np.array_equal(a, buf))
    assert hits > 200


def test_deblock_search_level_endpoints(oracle):
    """search_filter_level picks level 0 when the reconstruction already is
    the source, and a positive level when the reconstruction is a blocky
    version of a smooth source (deblocking removes the steps)."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        'zs', os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'zenav1-svt_amd', 'synth.py'))
    synth = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(synth)
    H, W = 64, 96
    geom = synth.partition_geometry(H // 4, W // 4, 11)
    yy, xx = np.mgrid[0:H, 0:W]
    smooth = (100 + 0.5 * xx + 0.25 * yy + 2 * np.sin(xx / 5.0)).astype(np.uint8)
    b, e0, be = oracle.deblock_search_level(smooth, smooth, 8, 0, 0, 0, geom, W // 4, H // 4, W, H)
    assert (b, e0, be) == (0, 0, 0)
    # 8x8 block means of a gentle ramp: small steps that the filters smooth
    blocky = smooth.astype(np.int32).copy()
    for y in range(0, H, 8):
        for x in range(0, W, 8):
            blocky[y:y + 8, x:x + 8] = blocky[y:y + 8, x:x + 8].mean()
    blocky = blocky.astype(np.uint8)
    for conv in (0, 1):
        b, e0, be = oracle.deblock_search_level(blocky, smooth, 8, 0, 0, 0, geom, W // 4, H // 4, W, H, conv)
        assert b > 0 and be < e0
