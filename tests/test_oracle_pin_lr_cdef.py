"""Pins of the oracle's self-guided statistics and CDEF filter taps by what
the mathematics and the AV1 specification fix, written independently of the
oracle's structure.

Self-guided (no counterpart in the reference, enc/restoration.rs:5-9; the
algorithm text is specs/08-loop-filters.md:371-420):
  * the 16 parameter sets and both lookup tables against the AV1
    specification's values (Sgr_Params; x_by_xplus1 from its integer
    expression ((z << 8) + z / 2) / (z + 1), 1 at z = 0, 256 from z = 255;
    one_by_x = round(4096 / n));
  * box sums, A / B, both filter passes (the r = 2 pass on every other row)
    and the five projection sums brute-forced in numpy over whole arrays
    (sliding windows on an edge-replicated plane), on units up to 12 x 12 at
    8 and 10 bits for all 16 sets.
CDEF (dsp/cdef.rs:931-1010; the reference's direction table test
:1226-1262):
  * single impulses at every primary / secondary tap position of every
    direction move the centre sample by exactly the tap weight's share
    (primary 4/2 at even strength, 3/3 at odd; secondary 2/1), and an
    impulse off the direction's taps leaves it unchanged.
"""
import ctypes

import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view

# AV1 specification Sgr_Params: r0, s0, r1, s1 (radius 0 = pass off)
SGR_PARAMS = [(2, 140, 1, 3236), (2, 112, 1, 2158), (2, 93, 1, 1618), (2, 80, 1, 1438), (2, 70, 1, 1295),
              (2, 58, 1, 1177), (2, 47, 1, 1079), (2, 37, 1, 996), (2, 30, 1, 925), (2, 25, 1, 863),
              (0, -1, 1, 2589), (0, -1, 1, 1618), (0, -1, 1, 1177), (0, -1, 1, 925), (2, 56, 0, -1),
              (2, 22, 0, -1)]
ONE_BY_X = [4096, 2048, 1365, 1024, 819, 683, 585, 512, 455, 410, 372, 341, 315, 293, 273, 256, 241, 228, 216,
            205, 195, 186, 178, 171, 164]


def x_by_xplus1(z):
    if z == 0:
        return 1
    if z >= 255:
        return 256
    return ((z << 8) + z // 2) // (z + 1)


def test_sgr_tables_match_specification(oracle):
    L = oracle.lrlib()
    L.or_sgr_table.argtypes = [ctypes.c_int, ctypes.c_int]
    for i, p in enumerate(SGR_PARAMS):
        assert tuple(L.or_sgr_param(i, k) for k in range(4)) == p, i
    assert [L.or_sgr_table(0, z) for z in range(256)] == [x_by_xplus1(z) for z in range(256)]
    assert [L.or_sgr_table(1, n - 1) for n in range(1, 26)] == ONE_BY_X
    # the integer expression is 256 z / (z + 1) rounded to nearest
    assert all(x_by_xplus1(z) == int(np.floor(256 * z / (z + 1) + 0.5)) for z in range(1, 255))
    assert ONE_BY_X == [int(np.floor(4096 / n + 0.5)) for n in range(1, 26)]


def _r2(v, n):
    return (v + (1 << (n - 1))) >> n if n else v


def _sgr_pass(pad, P, H, W, r, s, bd):
    """flt (H x W, units of 1/16) of one pass from an edge-padded plane slice
    `pad` whose [P:P+H, P:P+W] is the unit (P >= r + 1)."""
    n = (2 * r + 1) ** 2
    v = pad.astype(np.int64)
    # box sums centred on rows/cols -1 .. H / -1 .. W of the unit
    win = sliding_window_view(v, (2 * r + 1, 2 * r + 1))
    o = P - 1 - r
    win = win[o:o + H + 2, o:o + W + 2]
    bsum = win.sum(axis=(2, 3))
    bsq = (win * win).sum(axis=(2, 3))
    a = _r2(bsq, 2 * (bd - 8))
    b = _r2(bsum, bd - 8)
    p = np.maximum(a * n - b * b, 0)
    z = _r2(p * s, 20)
    A = np.vectorize(x_by_xplus1)(np.minimum(z, 255)).astype(np.int64)
    B = _r2((256 - A) * bsum * ONE_BY_X[n - 1], 12)
    d = v[P:P + H, P:P + W]

    def at(M, dy, dx):          # M indexed from -1: element (i + dy, j + dx)
        return M[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]

    if r == 1:
        wa = 4 * (at(A, 0, 0) + at(A, 0, -1) + at(A, 0, 1) + at(A, -1, 0) + at(A, 1, 0)) + \
            3 * (at(A, -1, -1) + at(A, -1, 1) + at(A, 1, -1) + at(A, 1, 1))
        wb = 4 * (at(B, 0, 0) + at(B, 0, -1) + at(B, 0, 1) + at(B, -1, 0) + at(B, 1, 0)) + \
            3 * (at(B, -1, -1) + at(B, -1, 1) + at(B, 1, -1) + at(B, 1, 1))
        return _r2(wa * d + wb, 9)
    # r = 2: A / B exist on odd rows only; even rows mix the rows above and
    # below, odd rows their own row
    even = np.zeros((H, W), bool)
    even[0::2] = True
    ea = 6 * (at(A, -1, 0) + at(A, 1, 0)) + 5 * (at(A, -1, -1) + at(A, -1, 1) + at(A, 1, -1) + at(A, 1, 1))
    eb = 6 * (at(B, -1, 0) + at(B, 1, 0)) + 5 * (at(B, -1, -1) + at(B, -1, 1) + at(B, 1, -1) + at(B, 1, 1))
    oa = 6 * at(A, 0, 0) + 5 * (at(A, 0, -1) + at(A, 0, 1))
    ob = 6 * at(B, 0, 0) + 5 * (at(B, 0, -1) + at(B, 0, 1))
    return np.where(even, _r2(ea * d + eb, 9), _r2(oa * d + ob, 8))


def _tdiv(a, n):
    q = abs(int(a)) // n
    return q if a >= 0 else -q


def sgr_stats_numpy(dgd, src, bd, lim, pset):
    hs, he, vs, ve = lim
    H, W = ve - vs, he - hs
    P = 4
    pad = np.pad(dgd.astype(np.int64), P, mode='edge')[vs:vs + H + 2 * P, hs:hs + W + 2 * P]
    r0, s0, r1, s1 = SGR_PARAMS[pset]
    u = dgd[vs:ve, hs:he].astype(np.int64) << 4
    sv = (src[vs:ve, hs:he].astype(np.int64) << 4) - u
    a = _sgr_pass(pad, P, H, W, r0, s0, bd) - u if r0 else np.zeros_like(u)
    b = _sgr_pass(pad, P, H, W, r1, s1, bd) - u if r1 else np.zeros_like(u)
    sums = [(a * a).sum(), (a * b).sum(), (b * b).sum(), (a * sv).sum(), (b * sv).sum()]
    return [_tdiv(x, H * W) for x in sums]


@pytest.mark.parametrize('bd', [8, 10])
@pytest.mark.parametrize('seed', range(3))
def test_sgr_stats_brute_force(oracle, bd, seed):
    rng = np.random.default_rng(100 * bd + seed)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    PH, PW = 20, 24
    kind = seed % 3
    if kind == 0:
        dgd = rng.integers(0, mx + 1, (PH, PW))
    elif kind == 1:
        yy, xx = np.mgrid[0:PH, 0:PW]
        dgd = np.clip(((xx * 7 + yy * 5) << (bd - 8)) + rng.integers(-3, 4, (PH, PW)), 0, mx)
    else:
        dgd = np.clip(rng.normal(mx / 2, mx / 40, (PH, PW)), 0, mx)
    dgd = dgd.astype(dt)
    src = np.clip(dgd.astype(np.int64) + rng.integers(-12 << (bd - 8), (12 << (bd - 8)) + 1, dgd.shape), 0,
                  mx).astype(dt)
    units = [(0, 12, 0, 12), (5, 17, 3, 15), (12, 24, 8, 20), (20, 24, 17, 20), (7, 8, 9, 10)]
    for lim in units:
        for pset in range(16):
            got = oracle.sgr_unit_stats(dgd, src, bd, lim, pset).tolist()
            assert got == sgr_stats_numpy(dgd, src, bd, lim, pset), (bd, seed, lim, pset)


# ---- CDEF taps by impulses --------------------------------------------------

SPEC_DIRS = [[(-1, 1), (-2, 2)], [(0, 1), (-1, 2)], [(0, 1), (0, 2)], [(0, 1), (1, 2)],
             [(1, 1), (2, 2)], [(1, 0), (2, 1)], [(1, 0), (2, 0)], [(1, 0), (2, -1)]]   # dsp/cdef.rs:1230-1239


def _constrain(d, t, damp):
    if t == 0:
        return 0
    sh = max(0, damp - (int(t).bit_length() - 1))
    v = min(abs(d), max(0, t - (abs(d) >> sh)))
    return v if d >= 0 else -v


def _centre(oracle, off, delta, pri, sec, dirn, pdamp, sdamp, cs):
    S = 144
    buf = np.full(S * 16, 400, np.uint16)
    c0 = 8 * S + 16
    if off is not None:
        buf[c0 + off[0] * S + off[1]] = 400 + delta
    out = np.zeros(1, np.int32)
    oracle.lflib().or_cdef_filter_block(out.ctypes.data, 1, buf[c0:].ctypes.data, pri, sec, dirn, pdamp, sdamp, 1,
                                        1, cs, 1)
    return int(out[0])


@pytest.mark.parametrize('dirn', range(8))
def test_cdef_impulse_taps(oracle, dirn):
    """10-bit samples (coeff_shift 2): flat 400 with one +160 impulse."""
    cs, delta, pdamp, sdamp = 2, 160, 12, 12
    for pri, weights in ((15 << 2, (3, 3)), (14 << 2, (4, 2))):
        cv = _constrain(delta, pri, pdamp)
        for k in range(2):
            for sgn in (1, -1):
                dy, dx = SPEC_DIRS[dirn][k]
                got = _centre(oracle, (sgn * dy, sgn * dx), delta, pri, 0, dirn, pdamp, sdamp, cs)
                assert got == 400 + ((8 + weights[k] * cv) >> 4), (dirn, pri, k, sgn, got)
    sec = 4 << 2
    cv = _constrain(delta, sec, sdamp)
    for side in (2, -2):
        d2 = (dirn + side) % 8
        for k, w in enumerate((2, 1)):
            for sgn in (1, -1):
                dy, dx = SPEC_DIRS[d2][k]
                got = _centre(oracle, (sgn * dy, sgn * dx), delta, 0, sec, dirn, pdamp, sdamp, cs)
                assert got == 400 + ((8 + w * cv) >> 4), (dirn, side, k, sgn, got)
    # an impulse on none of the direction's taps changes nothing
    taps = set()
    for d2, ks in ((dirn, (0, 1)), ((dirn + 2) % 8, (0, 1)), ((dirn - 2) % 8, (0, 1))):
        for k in ks:
            dy, dx = SPEC_DIRS[d2][k]
            taps |= {(dy, dx), (-dy, -dx)}
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if (dy, dx) != (0, 0) and (dy, dx) not in taps:
                assert _centre(oracle, (dy, dx), delta, 15 << 2, 4 << 2, dirn, pdamp, sdamp, cs) == 400


def test_cdef_constrain_reference_values(oracle):
    """The reference's hand-checked constrain points (dsp/cdef.rs:1270-1281)
    and the closed form over a grid."""
    L = oracle.lflib()
    assert L.or_cdef_constrain(1000, 0, 6) == 0
    assert L.or_cdef_constrain(5, 4, 4) == 3 and L.or_cdef_constrain(-5, 4, 4) == -3
    assert L.or_cdef_constrain(0x7f7f - 128, 15, 6) == 0 and L.or_cdef_constrain(0x4000 - 128, 15, 6) == 0
    assert L.or_cdef_constrain(2, 15, 3) == 2
    for d in range(-300, 301, 7):
        for t in (1, 2, 3, 4, 7, 15, 16, 60):
            for damp in (3, 4, 6, 8, 12):
                assert L.or_cdef_constrain(d, t, damp) == _constrain(d, t, damp)
