"""Pins of the oracle's Wiener unit filter (or_wiener_filter_plane), the
finer tap search (or_wiener_refine_unit) and the 10-bit SSE decomposition
(or_sse_region), following rust/crates/svtav1-dsp/src/restoration.rs:138
wiener_convolve_add_src, :1248 wiener_filter_stripe, :1288
loop_restoration_filter_unit (search path), :1921 its 10-bit twin, :1991
sse_region_hbd and rust/crates/svtav1-encoder/src/restoration.rs:513
finer_tile_search_wiener.  The reference's tests for these are differential
against the C library (absent here); what pins the oracle:

- zero taps are the identity (closed form of the two rounding stages);
- a numpy transcription of the reference's own loop structure (edge
  extension by 3, 64-row stripes offset by 8 >> ss_y, 64 >> ss_x
  processing units widened to a multiple of 16, the horizontal pass over
  h + 6 rows, the vertical over a temp with one zero row more) equals the
  oracle's per-pixel form on every unit, random signalable taps included;
- the 10-bit SSE truncates the right-strip partial sum to 32 bits (a case
  built to exceed 2^32) and not the plain sum;
- the hill climb, transcribed from the reference in Python over the numpy
  filter, gives the oracle's taps and SSE; the SSE never exceeds the
  starting one and the taps stay symmetric, within their signalable
  bounds, with the centre tap rebalanced.
"""
import numpy as np
import pytest

TAP_MIN = (3 - 8, -7 - 16, 15 - 32)
TAP_MAX = (3 - 1 + 8, -7 - 1 + 16, 15 - 1 + 32)


def random_filter(rng, win5):
    t0 = 0 if win5 else int(rng.integers(TAP_MIN[0], TAP_MAX[0] + 1))
    t1 = int(rng.integers(TAP_MIN[1], TAP_MAX[1] + 1))
    t2 = int(rng.integers(TAP_MIN[2], TAP_MAX[2] + 1))
    return np.array([t0, t1, t2, -2 * (t0 + t1 + t2), t2, t1, t0, 0], np.int16)


def convolve(src, r0, c0, hf, vf, w, h, bd):
    """wiener_convolve_add_src(_hbd) on a padded array at (r0, c0)."""
    ih = h + 6
    temp = np.zeros((ih + 1, w), np.int64)
    clamp = (1 << (bd + 1 + 7 - 3)) - 1
    for y in range(ih):
        row = src[r0 + y - 3].astype(np.int64)
        s = (row[c0 + 0:c0 + w] << 7) + (1 << (bd + 6))
        for k in range(8):
            s = s + row[c0 + k - 3:c0 + k - 3 + w] * int(hf[k])
        temp[y] = np.clip((s + 4) >> 3, 0, clamp)
    out = np.zeros((h, w), np.int64)
    for y in range(h):
        s = (temp[y + 3] << 7) - (1 << (bd + 10))
        for k in range(8):
            s = s + temp[y + k] * int(vf[k])
        out[y] = np.clip((s + 1024) >> 11, 0, (1 << bd) - 1)
    return out


def filter_plane_structured(dgd, bd, unit_size, ss_x, ss_y, types, filters):
    """The reference's unit loop: extend by 3 (and pad generously), every
    unit in raster order filtered stripe by stripe, processing unit by
    processing unit, into a padded destination."""
    ph, pw = dgd.shape
    B = 3
    P = 40  # room for the widened processing units and the 8th taps
    ext = np.zeros((ph + 2 * P, pw + 2 * P), np.int64)
    ext[P:P + ph, P:P + pw] = dgd
    ext[P:P + ph, P - B:P] = ext[P:P + ph, P:P + 1]
    ext[P:P + ph, P + pw:P + pw + B] = ext[P:P + ph, P + pw - 1:P + pw]
    for i in range(1, B + 1):
        ext[P - i, P - B:P + pw + B] = ext[P, P - B:P + pw + B]
        ext[P + ph - 1 + i, P - B:P + pw + B] = ext[P + ph - 1, P - B:P + pw + B]
    dst = np.zeros_like(ext)
    count = lambda s: max(1, (s + unit_size // 2) // unit_size)  # noqa: E731
    nh, nv = count(pw), count(ph)
    full = 64 >> ss_y
    roff = 8 >> ss_y
    puw = 64 >> ss_x
    for uy in range(nv):
        for ux in range(nh):
            u = uy * nh + ux
            e = unit_size * 3 // 2
            y0, x0 = uy * unit_size, ux * unit_size
            hh = ph - y0 if ph - y0 < e else unit_size
            ww = pw - x0 if pw - x0 < e else unit_size
            vs, ve = max(0, y0 - roff), y0 + hh
            if ve < ph:
                ve -= roff
            hs, he = x0, x0 + ww
            if not types[u]:
                dst[P + vs:P + ve, P + hs:P + he] = ext[P + vs:P + ve, P + hs:P + he]
                continue
            i = 0
            while i < ve - vs:
                v0 = vs + i
                stripe = (v0 + roff) // full
                h = min(full - (roff if stripe == 0 else 0), ve - v0)
                j = 0
                while j < he - hs:
                    w = min(puw, (he - hs - j + 15) & ~15)
                    dst[P + v0:P + v0 + h, P + hs + j:P + hs + j + w] = convolve(
                        ext, P + v0, P + hs + j, filters[u][0], filters[u][1], w, h, bd)
                    j += puw
                i += h
    return dst[P:P + ph, P:P + pw]


def plane(rng, h, w, bd):
    hi = (1 << bd) - 1
    p = np.clip(rng.normal(hi / 2, hi / 4, (h, w)), 0, hi)
    for _ in range(5):
        r, c = rng.integers(0, h - 4), rng.integers(0, w - 4)
        p[r:r + rng.integers(2, 30), c:c + rng.integers(2, 30)] = rng.integers(0, hi + 1)
    return p.astype(np.uint16 if bd > 8 else np.uint8)


def test_zero_taps_are_identity(oracle):
    rng = np.random.default_rng(1)
    for bd in (8, 10):
        d = plane(rng, 96, 80, bd)
        n = 2  # 1 x 2 units of 64
        out = oracle.wiener_filter_plane(d, bd, 64, 0, np.ones(n, np.uint8), np.zeros((n, 2, 8), np.int16))
        assert np.array_equal(out, d)


@pytest.mark.parametrize('bd,h,w,unit,ssx,ssy,win5,seed', [(8, 200, 232, 64, 0, 0, False, 1),
                                                            (10, 150, 136, 64, 1, 1, True, 2),
                                                            (8, 130, 300, 128, 0, 0, False, 3),
                                                            (10, 96, 96, 32, 1, 1, True, 4)])
def test_filter_plane_equals_reference_structure(oracle, bd, h, w, unit, ssx, ssy, win5, seed):
    rng = np.random.default_rng(seed)
    d = plane(rng, h, w, bd)
    nh, nv = max(1, (w + unit // 2) // unit), max(1, (h + unit // 2) // unit)
    n = nh * nv
    types = (rng.random(n) < 0.8).astype(np.uint8)
    filters = np.stack([np.stack([random_filter(rng, win5), random_filter(rng, win5)]) for _ in range(n)])
    filters[0] = [[7, -23, -17, 66, -17, -23, 7, 0], [-5, 8, 46, -98, 46, 8, -5, 0]]  # extreme: clamps engage
    types[0] = 1
    got = oracle.wiener_filter_plane(d, bd, unit, ssy, types, filters)
    want = filter_plane_structured(d.astype(np.int64), bd, unit, ssx, ssy, types, filters)
    assert np.array_equal(got.astype(np.int64), want), np.argwhere(got != want)[:5]


def test_sse_hbd_truncates_the_right_strip(oracle):
    L = oracle.lrlib()
    h, w = 400, 31              # right strip 15 columns x 400 rows
    a = np.zeros((h, w), np.uint16)
    b = np.full((h, w), 1023, np.uint16)
    strip = 15 * h * 1023 ** 2  # > 2^32
    blocks = 16 * 400 * 1023 ** 2
    assert strip > 2 ** 32
    got = L.or_sse_region(a.ctypes.data, b.ctypes.data, w, 10, 0, w, 0, h)
    assert got == blocks + (strip & 0xFFFFFFFF)
    assert L.or_sse_region(a.ctypes.data, b.ctypes.data, w, 10, 0, w, 0, h) != h * w * 1023 ** 2
    a8, b8 = np.zeros((h, w), np.uint8), np.full((h, w), 255, np.uint8)
    assert L.or_sse_region(a8.ctypes.data, b8.ctypes.data, w, 8, 0, w, 0, h) == h * w * 255 ** 2


def sse_hbd(a, b, bd):
    d = (a.astype(np.int64) - b.astype(np.int64)) ** 2
    if bd == 8:
        return int(d.sum())
    h, w = d.shape
    dw, dh = w % 16, h % 16
    t = 0
    if dw:
        t += int(d[:, w - dw:].sum()) & 0xFFFFFFFF
    if dh:
        t += int(d[h - dh:, :w - dw].sum()) & 0xFFFFFFFF
    t += int(d[:h - dh, :w - dw].sum())
    return t


def refine_py(dgd, src, bd, win, lim, hf, vf, max_one):
    """finer_tile_search_wiener transcribed over the numpy filter."""
    hs, he, vs, ve = lim
    ph, pw = dgd.shape
    pad = np.pad(dgd.astype(np.int64), 8, mode='edge')

    def err_of(hf_, vf_):
        out = convolve(pad, 8 + vs, 8 + hs, hf_, vf_, he - hs, ve - vs, bd)
        return sse_hbd(src[vs:ve, hs:he], out, bd)

    hf, vf = list(map(int, hf)), list(map(int, vf))
    err = err_of(hf, vf)
    start, end = 4, (4 if max_one else 1)
    s = start
    while s >= end:
        for f in (hf, vf):
            for p in range((7 - win) >> 1, 3):
                skip = False
                while True:
                    if f[p] - s >= TAP_MIN[p]:
                        f[p] -= s; f[6 - p] -= s; f[3] += 2 * s
                        e2 = err_of(hf, vf)
                        if e2 > err:
                            f[p] += s; f[6 - p] += s; f[3] -= 2 * s
                        else:
                            err = e2
                            skip = True
                            if s == start and not max_one:
                                continue
                    break
                if skip:
                    break
                while True:
                    if f[p] + s <= TAP_MAX[p]:
                        f[p] += s; f[6 - p] += s; f[3] -= 2 * s
                        e2 = err_of(hf, vf)
                        if e2 > err:
                            f[p] -= s; f[6 - p] -= s; f[3] += 2 * s
                        else:
                            err = e2
                            if s == start and not max_one:
                                continue
                    break
        s >>= 1
    return err, hf, vf


@pytest.mark.parametrize('bd,win,max_one,seed', [(8, 7, False, 11), (10, 5, False, 12), (8, 7, True, 13)])
def test_refine_matches_transcription_and_invariants(oracle, bd, win, max_one, seed):
    rng = np.random.default_rng(seed)
    src = plane(rng, 48, 40, bd)
    hi = (1 << bd) - 1
    dgd = np.clip(src.astype(np.int64) + rng.integers(-hi // 12, hi // 12 + 1, src.shape), 0, hi).astype(src.dtype)
    lim = (0, 40, 0, 40)
    hf0, vf0 = random_filter(rng, win == 5), random_filter(rng, win == 5)
    e0 = oracle.wiener_refine_unit(dgd, src, bd, win, lim, hf0, vf0, use_refinement=False)[0]
    e, hf, vf = oracle.wiener_refine_unit(dgd, src, bd, win, lim, hf0, vf0, max_one_step=max_one)
    pe, phf, pvf = refine_py(dgd, src, bd, win, lim, hf0, vf0, max_one)
    assert (e, list(hf), list(vf)) == (pe, phf, pvf)
    assert e <= e0
    for f in (hf, vf):
        assert f[7] == 0 and all(f[k] == f[6 - k] for k in range(3)) and f[3] == -2 * (f[0] + f[1] + f[2])
        assert all(TAP_MIN[k] <= f[k] <= TAP_MAX[k] for k in range(3))
        if win == 5:
            assert f[0] == 0
