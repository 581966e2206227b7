"""Pins of the oracle's CDEF frame pass (or_cdef_apply_frame, following
apply_cdef_frame / apply_cdef_frame_hbd, rust/crates/svtav1-encoder/src/
cdef.rs:343 / :556).  The reference's tests hold no frame-level vectors for
this pass; its parts are pinned elsewhere (cdef_find_dir and the block
filter's taps by the reference's direction table and hand-checked constrain
values, tests/test_oracle_pin_lr_cdef.py).  What the frame composition adds
is pinned here by what it must satisfy:

- zero strengths and an all-skip map leave every plane unchanged;
- a strength in one filter block changes pixels of that block only (and
  its chroma counterpart), and never a skipped 8x8;
- an 8x8 with one non-skip mi is filtered (the skip rule needs all four);
- a per-block composition written here from the reference's text (padded
  fb source, CDEF_VERY_LARGE outside the plane, strength decode with
  secondary 3 -> 4, the luma direction zeroed at primary 0, chroma damping
  one lower, both shifted by bd - 8) through the pinned block functions
  equals the frame pass;
- the evidence counters count 64 / 16 pixels per filtered luma / chroma
  block and the pixels that differ.
"""
import numpy as np
import pytest

LARGE = 0x7F7F
VB, HB, S = 3, 8, 144


def frame(rng, h, w, bd):
    hi = (1 << bd) - 1
    y = np.clip(rng.normal(hi / 2, hi / 6, (h, w)), 0, hi).astype(np.uint16 if bd > 8 else np.uint8)
    # blocky regions so directions and clamps vary
    for _ in range(6):
        r, c = rng.integers(0, h - 8), rng.integers(0, w - 8)
        y[r:r + rng.integers(4, 24), c:c + rng.integers(4, 24)] = rng.integers(0, hi + 1)
    u = np.clip(rng.normal(hi / 2, hi / 8, (h // 2, w // 2)), 0, hi).astype(y.dtype)
    v = np.clip(rng.normal(hi / 2, hi / 8, (h // 2, w // 2)), 0, hi).astype(y.dtype)
    return y, u, v


def pad_src(plane, y0, x0, vsize, hsize):
    ph, pw = plane.shape
    buf = np.full(S * (64 + 2 * VB), LARGE, np.uint16)
    for r in range(vsize + 2 * VB):
        gy = y0 + r - VB
        if not 0 <= gy < ph:
            continue
        for c in range(hsize + 2 * HB):
            gx = x0 + c - HB
            if 0 <= gx < pw:
                buf[r * S + c] = plane[gy, gx]
    return buf


def compose(oracle, y, u, v, bd, damping, fb_str, skip):
    """The frame pass block by block through the block functions."""
    L = oracle.lflib()
    h, w = y.shape
    cs = bd - 8
    nhfb, nvfb = (w + 63) // 64, (h + 63) // 64
    oy, ou, ov = y.astype(np.int32), u.astype(np.int32), v.astype(np.int32)
    mr, mc = skip.shape
    for fbr in range(nvfb):
        for fbc in range(nhfb):
            vs, hs = min(64, h - fbr * 64), min(64, w - fbc * 64)
            act = {}
            for by in range(vs // 8):
                for bx in range(hs // 8):
                    m = skip[fbr * 16 + by * 2:min(fbr * 16 + by * 2 + 2, mr), fbc * 16 + bx * 2:min(fbc * 16 + bx * 2 + 2, mc)]
                    if not m.all():
                        act[by, bx] = True
            if not act:
                continue
            sy, suv = fb_str[fbr * nhfb + fbc]
            py, sy2 = (sy // 4) << cs, ((sy % 4) + (sy % 4 == 3)) << cs
            pu, su2 = (suv // 4) << cs, ((suv % 4) + (suv % 4 == 3)) << cs
            pad = pad_src(y, fbr * 64, fbc * 64, vs, hs)
            base = VB * S + HB
            dirs = {}
            for (by, bx) in act:
                o = base + by * 8 * S + bx * 8
                blk = np.stack([pad[o + i * S:o + i * S + 8] for i in range(8)])
                dirs[by, bx] = oracle.cdef_find_dir(blk, cs)
            if py or sy2:
                for (by, bx) in act:
                    d, var = dirs[by, bx]
                    t = L.or_cdef_adjust_strength(py, var)
                    if t == 0 and sy2 == 0:
                        continue
                    out = np.zeros(64, np.int32)
                    L.or_cdef_filter_block(out.ctypes.data, 8, pad[base + by * 8 * S + bx * 8:].ctypes.data, t, sy2,
                                           d if py else 0, damping + cs, damping + cs, 8, 8, cs, 1)
                    oy[fbr * 64 + by * 8:fbr * 64 + by * 8 + 8, fbc * 64 + bx * 8:fbc * 64 + bx * 8 + 8] = out.reshape(8, 8)
            if pu or su2:
                for pl, (src, dst) in enumerate(((u, ou), (v, ov))):
                    padc = pad_src(src, fbr * 32, fbc * 32, vs // 2, hs // 2)
                    for (by, bx) in act:
                        out = np.zeros(16, np.int32)
                        L.or_cdef_filter_block(out.ctypes.data, 4, padc[base + by * 4 * S + bx * 4:].ctypes.data, pu,
                                               su2, dirs[by, bx][0] if pu else 0, damping + cs - 1, damping + cs - 1,
                                               4, 4, cs, 1)
                        dst[fbr * 32 + by * 4:fbr * 32 + by * 4 + 4, fbc * 32 + bx * 4:fbc * 32 + bx * 4 + 4] = \
                            out.reshape(4, 4)
    return oy, ou, ov


@pytest.mark.parametrize('bd', [8, 10])
def test_cdef_apply_identities(oracle, bd):
    rng = np.random.default_rng(60 + bd)
    y, u, v = frame(rng, 128, 192, bd)
    nfb = 2 * 3
    oy, ou, ov, st = oracle.cdef_apply_frame(y, u, v, bd, 5, np.zeros((nfb, 2), np.uint8))
    assert np.array_equal(oy, y) and np.array_equal(ou, u) and np.array_equal(ov, v) and st.tolist() == [0, 0]
    skip = np.ones((32, 48), np.uint8)
    strong = np.full((nfb, 2), 63, np.uint8)
    oy, ou, ov, st = oracle.cdef_apply_frame(y, u, v, bd, 5, strong, skip)
    assert np.array_equal(oy, y) and np.array_equal(ou, u) and np.array_equal(ov, v) and st.tolist() == [0, 0]


def test_cdef_apply_locality_and_skip_rule(oracle):
    rng = np.random.default_rng(7)
    y, u, v = frame(rng, 128, 128, 8)
    fs = np.zeros((4, 2), np.uint8)
    fs[3] = (4 * 9 + 2, 4 * 5 + 3)  # the bottom-right fb only
    skip = np.ones((32, 32), np.uint8)
    skip[16:32, 16:32] = 0          # that fb not skipped ...
    skip[20:22, 20:22] = 1          # ... but for one whole 8x8 (mi 20..21)
    skip[24, 24] = 1                # one mi of another 8x8: still filtered
    oy, ou, ov, st = oracle.cdef_apply_frame(y, u, v, 8, 4, fs, skip)
    dy = oy.astype(int) != y
    assert dy.any() and not dy[:64].any() and not dy[:, :64].any()
    assert not dy[80:88, 80:88].any()                  # the all-skip 8x8
    assert dy[96:104, 96:104].any()                    # one non-skip mi of four
    du = (ou.astype(int) != u) | (ov.astype(int) != v)
    assert du.any() and not du[:32].any() and not du[:, :32].any() and not du[40:44, 40:44].any()
    # counters: 63 filtered luma 8x8s (the adjusted primary or the secondary
    # is nonzero: the secondary is 2 here), 63 chroma 4x4s in each of two planes
    assert int(st[0]) == 63 * 64 + 2 * 63 * 16
    assert int(st[1]) == int(dy.sum() + (ou.astype(int) != u).sum() + (ov.astype(int) != v).sum())


@pytest.mark.parametrize('bd,h,w,seed', [(8, 136, 200, 1), (10, 96, 128, 2), (8, 64, 72, 3)])
def test_cdef_apply_equals_block_composition(oracle, bd, h, w, seed):
    rng = np.random.default_rng(seed)
    y, u, v = frame(rng, h, w, bd)
    nfb = ((w + 63) // 64) * ((h + 63) // 64)
    fs = rng.integers(0, 64, (nfb, 2)).astype(np.uint8)
    fs[0] = (0, 3)       # primary 0: direction 0, secondary 4
    mr, mc = (h + 3) // 4, (w + 3) // 4
    skip = (rng.random((mr, mc)) < 0.35).astype(np.uint8)
    damping = int(rng.integers(3, 7))
    oy, ou, ov, _ = oracle.cdef_apply_frame(y, u, v, bd, damping, fs, skip)
    ry, ru, rv = compose(oracle, y, u, v, bd, damping, fs, skip)
    assert np.array_equal(oy.astype(np.int32), ry)
    assert np.array_equal(ou.astype(np.int32), ru) and np.array_equal(ov.astype(np.int32), rv)
