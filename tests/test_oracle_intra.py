"""Pin the oracle's intra predictors: the reference's known answers
(golden_parity.rs:488-930), closed forms of the 45/135-degree and V/H
directions, constant-edge invariance of every filter/upsample path, and
filter-intra / CfL identities (CPU)."""
import numpy as np
import pytest


def job(oracle, x, y, w, h, mode, delta=0, fi=5, ef=0, ft=0, nt=None, ntr=0, nl=None, nbl=0):
    j = np.zeros(1, oracle.INTRA_JOB)
    j['x'], j['y'], j['w'], j['h'], j['mode'], j['delta'] = x, y, w, h, mode, delta
    j['fi_mode'], j['edge_filter'], j['filt_type'] = fi, ef, ft
    j['n_top_px'] = w if nt is None else nt
    j['n_left_px'] = h if nl is None else nl
    j['n_topright_px'], j['n_bottomleft_px'] = ntr, nbl
    return j


def plane_with_edges(above, left, tl, dtype=np.uint8, size=80):
    """A plane whose block at (8, 8) has the given neighbours."""
    p = np.zeros((size, size), dtype)  # block at (8, 8)
    p[7, 8:8 + len(above)] = above
    p[8:8 + len(left), 7] = left
    p[7, 7] = tl
    return p


def test_reference_known_answers(oracle):
    # DC of above [10,20,30,40] + left [80,70,60,50] = 45 (golden_parity.rs:504-518)
    p = plane_with_edges([10, 20, 30, 40], [80, 70, 60, 50], 0)
    assert np.all(oracle.intra_predict(p, 8, job(oracle, 8, 8, 4, 4, 0)) == 45)
    # V / H copy exactly (:520-547)
    p = plane_with_edges([10, 20, 30, 40], [10, 20, 30, 40], 0)
    assert np.all(oracle.intra_predict(p, 8, job(oracle, 8, 8, 4, 4, 1)) == [10, 20, 30, 40])
    assert np.all(oracle.intra_predict(p, 8, job(oracle, 8, 8, 4, 4, 2)).T == [10, 20, 30, 40])
    # Paeth (:549-563): tl 50, above [10..40], left [60..90] -> (0,0)=10, (0,3)=50
    p = plane_with_edges([10, 20, 30, 40], [60, 70, 80, 90], 50)
    d = oracle.intra_predict(p, 8, job(oracle, 8, 8, 4, 4, 12))
    assert d[0, 0] == 10 and d[0, 3] == 50


def test_paeth_tie_orders_agree(oracle):
    """The 8-bit Paeth breaks ties top-first (intra_pred.rs:282), the 10-bit
    one left-first (hbd.rs:203).  A tie between top and left only occurs when
    both samples are equal, so the two orders must give identical output:
    checked over 64 x 64 (top, left) pairs for 8 top-left values."""
    rng = np.random.default_rng(0)
    for tl in (0, 1, 64, 127, 128, 200, 254, 255):
        above = rng.permutation(256)[:64]
        left = rng.permutation(256)[:64]
        p8 = plane_with_edges(above, left, tl)
        p10 = plane_with_edges(above, left, tl, np.uint16)
        a = oracle.intra_predict(p8, 8, job(oracle, 8, 8, 64, 64, 12))
        b = oracle.intra_predict(p10, 10, job(oracle, 8, 8, 64, 64, 12))
        assert np.array_equal(a, b)
        # each output is one of its three candidates
        assert np.all((a == above[None, :]) | (a == left[:, None]) | (a == tl))


@pytest.mark.parametrize('bd', [8, 10])
def test_directional_closed_forms(oracle, bd):
    """D45 (dx = 64): pred[r][c] = above[r+c+1]; D135 (dx = dy = 64):
    pred = above[c-r-1] / top-left / left[r-c-1]; no edge filter."""
    dt = np.uint8 if bd == 8 else np.uint16
    rng = np.random.default_rng(bd)
    w = h = 8
    above = rng.integers(0, 1 << bd, 2 * w)
    left = rng.integers(0, 1 << bd, 2 * h)
    tl = int(rng.integers(0, 1 << bd))
    p = plane_with_edges(above, left, tl, dt)
    d45 = oracle.intra_predict(p, bd, job(oracle, 8, 8, w, h, 3, ntr=w))
    for r in range(h):
        for c in range(w):
            assert d45[r, c] == above[min(r + c + 1, w + h - 1)]
    d135 = oracle.intra_predict(p, bd, job(oracle, 8, 8, w, h, 4))
    for r in range(h):
        for c in range(w):
            exp = above[c - r - 1] if c > r else (tl if c == r else left[r - c - 1])
            assert d135[r, c] == exp


@pytest.mark.parametrize('bd', [8, 10])
def test_constant_edges_predict_constant(oracle, bd):
    """Every directional angle, with edge filtering and upsampling, and every
    filter-intra mode maps a constant neighbourhood to the same constant."""
    dt = np.uint8 if bd == 8 else np.uint16
    v = 77 if bd == 8 else 601
    p = np.full((96, 96), v, dt)
    for (w, h) in ((4, 4), (8, 8), (4, 16), (16, 8), (32, 32), (64, 16)):
        for mode in range(13):
            for delta in (-3, -1, 0, 2, 3):
                for ef, ft in ((0, 0), (1, 0), (1, 1)):
                    d = oracle.intra_predict(p, bd, job(oracle, 16, 16, w, h, mode, delta, 5, ef, ft, ntr=w, nbl=h))
                    assert np.all(d == v), (w, h, mode, delta, ef, ft)
        if w <= 32 and h <= 32:
            for fi in range(5):
                d = oracle.intra_predict(p, bd, job(oracle, 16, 16, w, h, 0, 0, fi))
                assert np.all(d == v), (w, h, fi)


def test_unavailable_fallbacks(oracle):
    """No neighbours: DC = 128 << (bd-8); V uses base-1, H base+1 (partition.rs:330-370)."""
    for bd, base in ((8, 128), (10, 512)):
        p = np.zeros((16, 16), np.uint8 if bd == 8 else np.uint16)
        j0 = job(oracle, 0, 0, 4, 4, 0, nt=0, nl=0)
        assert np.all(oracle.intra_predict(p, bd, j0) == base)
        assert np.all(oracle.intra_predict(p, bd, job(oracle, 0, 0, 4, 4, 1, nt=0, nl=0)) == base - 1)
        assert np.all(oracle.intra_predict(p, bd, job(oracle, 0, 0, 4, 4, 2, nt=0, nl=0)) == base + 1)


def test_edge_filter_tables(oracle):
    L = oracle.lib()
    L.or_edge_filter_strength.argtypes = [oracle.ctypes.c_int] * 4
    L.or_use_upsample.argtypes = [oracle.ctypes.c_int] * 4
    # strength grows with block size and angle distance; >= 3 for large blocks
    assert L.or_edge_filter_strength(32, 32, 3, 0) == 3
    assert L.or_edge_filter_strength(4, 4, 3, 0) == 0
    assert L.or_edge_filter_strength(8, 16, 16, 0) == 2
    assert L.or_use_upsample(4, 4, 3, 0) == 1 and L.or_use_upsample(16, 16, 3, 0) == 0
    assert L.or_use_upsample(4, 4, 40, 0) == 0 and L.or_use_upsample(4, 4, 0, 0) == 0


@pytest.mark.parametrize('bd', [8, 10])
def test_cfl_identities(oracle, bd):
    rng = np.random.default_rng(bd)
    dt = np.uint8 if bd == 8 else np.uint16
    luma = rng.integers(0, 1 << bd, (32, 32)).astype(dt)
    dc = rng.integers(0, 1 << bd, (8, 8)).astype(np.int32)
    assert np.array_equal(oracle.cfl_predict(luma, bd, 8, 8, dc, 0), dc)           # alpha 0
    flat = np.full((32, 32), 300 if bd == 10 else 100, dt)
    assert np.array_equal(oracle.cfl_predict(flat, bd, 8, 8, dc, 16), dc)          # no AC
    up = oracle.cfl_predict(luma, bd, 8, 8, np.full((8, 8), 1 << (bd - 1)), 8)
    dn = oracle.cfl_predict(luma, bd, 8, 8, np.full((8, 8), 1 << (bd - 1)), -8)
    mid = 1 << (bd - 1)
    assert np.all(np.abs((up - mid) + (dn - mid)) <= 1)                           # antisymmetric in alpha
