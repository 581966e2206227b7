"""Pin the oracle's loop-restoration statistics: the reference's
instrumented-C LR captures (per-unit SSE on the C post-CDEF recon dumps),
unit geometry, and the covariance definition of the Wiener statistics;
self-guided statistics are spec-derived (parity unpinned) and checked only
by invariants (CPU)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def gradient_luma(w, h):
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    return (((r * 255) // h) ^ ((c * 3) & 0x3f)).astype(np.uint8)


def test_sse_region_matches_c_capture(oracle):
    d = json.load(open(os.path.join(HERE, 'golden', 'lr', 'README.json')))
    L = oracle.lrlib()
    for cell in d['cells']:
        dgd = np.fromfile(os.path.join(HERE, 'golden', 'lr', cell['dgd']), np.uint8).reshape(cell['h'], cell['w'])
        src = gradient_luma(cell['w'], cell['h'])
        hs, he, vs, ve = cell['limits']
        got = L.or_sse_region(src.ctypes.data, dgd.ctypes.data, cell['w'], 8, hs, he, vs, ve)
        assert got == cell['sse_none']


def test_unit_geometry(oracle):
    L = oracle.lrlib()
    assert L.or_count_units(128, 4096) == 32 and L.or_count_units(256, 64) == 1 and L.or_count_units(64, 95) == 1
    assert L.or_count_units(64, 96) == 2
    # 64x64 plane, unit 256: one unit covering everything (capture lim=[0,64,0,64])
    assert oracle.unit_limits(256, 64, 64, 0, 0, 0) == (0, 64, 0, 64)
    # interior rows shifted up by RESTORATION_UNIT_OFFSET (8 luma, 4 chroma)
    assert oracle.unit_limits(128, 512, 512, 0, 1, 1) == (128, 256, 120, 248)
    assert oracle.unit_limits(64, 256, 256, 1, 0, 3) == (0, 64, 188, 256)
    # last unit absorbs a remainder under 1.5 units
    assert oracle.unit_limits(128, 300, 300, 0, 1, 1) == (128, 300, 120, 300)


def _patch_stats(dgd, src, win, lim, bd):
    """Covariance form: Y = matrix of mean-removed windows (column-major
    window order), M = Y^T x, H = Y^T Y, computed with numpy."""
    hs, he, vs, ve = lim
    ph, pw = dgd.shape
    half = win // 2
    pad = np.pad(dgd.astype(np.int64), half, mode='edge')
    avg = int(dgd[vs:ve, hs:he].astype(np.int64).sum() // ((ve - vs) * (he - hs)))
    cols = []
    for k in range(-half, half + 1):      # column offset (outer)
        for l in range(-half, half + 1):  # row offset (inner)
            cols.append(pad[vs + l + half:ve + l + half, hs + k + half:he + k + half].ravel() - avg)
    Y = np.stack(cols, 1)
    x = src[vs:ve, hs:he].astype(np.int64).ravel() - avg
    M, H = Y.T @ x, Y.T @ Y
    if bd == 10:
        M = np.trunc(M / 4).astype(np.int64)
        H = np.trunc(H / 4).astype(np.int64)
    return M, H


@pytest.mark.parametrize('bd,win', [(8, 7), (8, 5), (10, 7), (10, 5)])
def test_wiener_stats_covariance(oracle, bd, win):
    rng = np.random.default_rng(bd + win)
    dt = np.uint8 if bd == 8 else np.uint16
    dgd = rng.integers(0, 1 << bd, (40, 56)).astype(dt)
    src = np.clip(dgd.astype(np.int64) + rng.integers(-9, 10, dgd.shape), 0, (1 << bd) - 1).astype(dt)
    for lim in ((0, 56, 0, 40), (5, 31, 3, 29), (40, 56, 20, 40)):
        M, H = oracle.wiener_stats(win, dgd, src, bd, lim)
        rM, rH = _patch_stats(dgd, src, win, lim, bd)
        assert np.array_equal(M, rM) and np.array_equal(H, rH)


def test_sgr_invariants(oracle):
    """Constant neighbourhood: both passes reproduce the sample (x16) up to the
    B-rounding slack, so the projection statistics vanish; a sample-exact
    source gives C == 0 whatever the filter output."""
    L = oracle.lrlib()
    assert [L.or_sgr_param(0, k) for k in range(4)] == [2, 140, 1, 3236]
    flat = np.full((48, 48), 120, np.uint8)
    for pset in range(16):
        st = oracle.sgr_unit_stats(flat, flat, 8, (8, 40, 8, 40), pset)
        assert np.all(np.abs(st) <= 4), (pset, st)
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (48, 48)).astype(np.uint8)
    for pset in (0, 10, 14):
        st = oracle.sgr_unit_stats(img, img, 8, (0, 48, 0, 48), pset)
        assert st[3] == 0 and st[4] == 0 and st[0] >= 0 and st[2] >= 0
