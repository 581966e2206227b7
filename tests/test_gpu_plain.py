"""GPU parity of the plain per-call table's pixel, prediction and
loop-filter entries (ABI 2) against the oracle, bit-exact: residual, recon,
SSE (8 / 16-bit), the non-directional predictors and filter intra from
given edges, the edged directional predictor, CfL, every deblocking filter
(both directions, 8 / 10-bit), cdef_find_dir, cdef_filter_block and
compute_stats.  Each call is one staged launch on the calling thread's
stream, as a per-call FFI binding uses it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = ctypes.POINTER


def _ptr(a, t=ctypes.c_void_p):
    return a.ctypes.data_as(t) if t is not ctypes.c_void_p else ctypes.c_void_p(a.ctypes.data)


@pytest.fixture(scope='module')
def tab(gpu):
    t = gpu.dsp_table()
    assert t.abi_version == 3
    return t


@pytest.mark.parametrize('bd', [8, 10])
def test_residual_recon_sse(tab, bd):
    rng = np.random.default_rng(bd)
    dt = np.uint8 if bd == 8 else np.uint16
    mx = (1 << bd) - 1
    for (w, h) in [(4, 4), (8, 16), (64, 32), (128, 128), (3, 5)]:
        ss, ps = w + 7, w + 3
        src = rng.integers(0, mx + 1, (h, ss)).astype(dt)
        pred = rng.integers(0, mx + 1, (h, ps)).astype(dt)
        out = np.zeros((h, w), np.int32)
        f = tab.residual_i32 if bd == 8 else tab.residual_i32_u16
        assert f(_ptr(src), ss, _ptr(pred), ps, w, h, out.ctypes.data_as(P(ctypes.c_int32))) == 0
        assert np.array_equal(out, src[:, :w].astype(np.int32) - pred[:, :w].astype(np.int32))
        inv = rng.integers(-(2 * mx), 2 * mx, (h, w)).astype(np.int32)
        rec = np.zeros((h, w), dt)
        if bd == 8:
            rc = tab.recon_add_clamp(_ptr(pred), ps, inv.ctypes.data_as(P(ctypes.c_int32)), w, h, _ptr(rec))
        else:
            rc = tab.recon_add_clamp_u16(_ptr(pred), ps, inv.ctypes.data_as(P(ctypes.c_int32)), w, h, _ptr(rec), bd)
        assert rc == 0
        assert np.array_equal(rec, np.clip(pred[:, :w].astype(np.int64) + inv, 0, mx).astype(dt))
        o = ctypes.c_uint64()
        f = tab.sse_u8 if bd == 8 else tab.sse_u16
        assert f(_ptr(src), ss, _ptr(pred), ps, w, h, ctypes.byref(o)) == 0
        d = src[:, :w].astype(np.int64) - pred[:, :w].astype(np.int64)
        assert o.value == int((d * d).sum())


def _plane(rng, bd, n=96):
    mx = (1 << bd) - 1
    yy, xx = np.mgrid[0:n, 0:n]
    p = (np.sin(xx / 5.0) * np.cos(yy / 7.0) * 0.4 + 0.5) * mx + rng.normal(0, 12 << (bd - 8), (n, n))
    return np.clip(p, 0, mx).astype(np.uint8 if bd == 8 else np.uint16)


@pytest.mark.parametrize('bd', [8, 10])
def test_nondirectional_and_filter_intra_from_edges(tab, oracle, bd):
    rng = np.random.default_rng(20 + bd)
    plane = _plane(rng, bd)
    x = y = 16
    for (w, h) in [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 16), (16, 4), (8, 32), (64, 16)]:
        above = np.ascontiguousarray(plane[y - 1, x:x + w])
        left = np.ascontiguousarray(plane[y:y + h, x - 1])
        tl = int(plane[y - 1, x - 1])
        for mode in (0, 1, 2, 9, 10, 11, 12):
            for ha, hl in ([(1, 1), (1, 0), (0, 1), (0, 0)] if mode == 0 else [(1, 1)]):
                job = np.zeros(1, oracle.INTRA_JOB)
                job[0] = (x, y, w, h, mode, 0, 5, 0, 0, 0, w if ha else 0, 0, h if hl else 0, 0, 0, 0)
                ref = oracle.intra_predict(plane, bd, job)
                got = np.zeros((h, w), plane.dtype)
                assert tab.intra_pred(mode, _ptr(got), w, _ptr(above), _ptr(left), tl, w, h, ha, hl, bd) == 0
                assert np.array_equal(got, ref), (w, h, mode, ha, hl)
        if w <= 32 and h <= 32:
            above_tl = np.ascontiguousarray(plane[y - 1, x - 1:x + w])
            for fi in range(5):
                job = np.zeros(1, oracle.INTRA_JOB)
                job[0] = (x, y, w, h, 0, 0, fi, 0, 0, 0, w, 0, h, 0, 0, 0)
                ref = oracle.intra_predict(plane, bd, job)
                got = np.zeros((h, w), plane.dtype)
                assert tab.predict_filter_intra(_ptr(got), w, _ptr(above_tl), _ptr(left), w, h, fi, bd) == 0
                assert np.array_equal(got, ref), (w, h, fi)


@pytest.mark.parametrize('bd', [8, 10])
def test_dr_predictor_edged(tab, oracle, bd):
    rng = np.random.default_rng(30 + bd)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    angles = [36, 45, 54, 58, 76, 87, 90, 93, 104, 113, 122, 126, 135, 144, 148, 157, 166, 177, 180, 183, 194,
              203, 212]
    for (w, h) in [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (16, 4), (8, 32), (64, 16)]:
        for a in angles:
            for ua, ul in ([(0, 0), (1, 0), (0, 1), (1, 1)] if w + h <= 16 else [(0, 0)]):
                above = rng.integers(0, mx + 1, 160).astype(dt)
                left = rng.integers(0, mx + 1, 160).astype(dt)
                left[15] = above[15]
                ref = oracle.dr_predict_edges(above.astype(np.int32), left.astype(np.int32), w, h, a, ua, ul, bd)
                got = np.zeros((h, w), dt)
                assert tab.dr_predictor_edged(_ptr(got), w, _ptr(above), _ptr(left), 160, 16, ua, ul, w, h, a,
                                              bd) == 0
                assert np.array_equal(got, ref.astype(dt)), (w, h, a, ua, ul)
        # a buffer whose origin is not 16
        above = rng.integers(0, mx + 1, 200).astype(dt)
        left = rng.integers(0, mx + 1, 200).astype(dt)
        left[29] = above[29]
        ref = oracle.dr_predict_edges(above[14:174].astype(np.int32), left[14:174].astype(np.int32), w, h, 122,
                                      0, 0, bd)
        got = np.zeros((h, w), dt)
        assert tab.dr_predictor_edged(_ptr(got), w, _ptr(above), _ptr(left), 200, 30, 0, 0, w, h, 122, bd) == 0
        assert np.array_equal(got, ref.astype(dt))
        # angles the encoder never forms and upsampling on large units are rejected
        assert tab.dr_predictor_edged(_ptr(got), w, _ptr(above), _ptr(left), 200, 30, 0, 0, w, h, 100, bd) == -22
        if w + h > 16:
            assert tab.dr_predictor_edged(_ptr(got), w, _ptr(above), _ptr(left), 200, 30, 1, 0, w, h, 45, bd) == -22


@pytest.mark.parametrize('bd', [8, 10])
def test_cfl_predict(tab, oracle, bd):
    rng = np.random.default_rng(40 + bd)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    for (cw, ch) in [(4, 4), (8, 8), (16, 16), (32, 32), (4, 16), (32, 8)]:
        luma = rng.integers(0, mx + 1, (2 * ch, 2 * cw)).astype(dt)
        dc = rng.integers(0, mx + 1, (ch, cw)).astype(dt)
        for alpha in (-16, -5, 0, 3, 16):
            ref = oracle.cfl_predict(luma, bd, cw, ch, dc.astype(np.int32), alpha)
            s = luma.astype(np.int32)
            q3 = ((s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2]) * 2).astype(np.int16)
            l2 = int(np.log2(cw * ch))
            avg = (int(q3.astype(np.int32).sum()) + ((1 << l2) >> 1)) >> l2
            ac = np.zeros((ch, 32), np.int16)
            ac[:, :cw] = (q3.astype(np.int32) - avg).astype(np.int16)
            got = np.zeros((ch, cw), dt)
            rc = tab.cfl_predict(ac.ctypes.data_as(P(ctypes.c_int16)), _ptr(dc), cw, _ptr(got), cw, alpha, cw, ch, bd)
            assert rc == 0
            assert np.array_equal(got, ref.astype(dt)), (cw, ch, alpha)


@pytest.mark.parametrize('bd', [8, 10])
def test_lpf_all_filters(tab, oracle, bd):
    rng = np.random.default_rng(50 + bd)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    L = oracle.lflib()
    n = 32
    for taps in (4, 6, 8, 14):
        for vertical in (0, 1):
            for level, sharp in ((8, 0), (30, 0), (63, 0), (40, 3), (20, 7)):
                for trial in range(4):
                    base = rng.integers(0, mx + 1)
                    buf = np.clip(base + rng.integers(-(3 << (bd - 8)), 3 << (bd - 8), (n, n))
                                  + (np.arange(n)[None, :] >= 16) * rng.integers(-20, 21) * (1 << (bd - 8)), 0,
                                  mx).astype(dt)
                    if vertical:
                        off = 10 * n + 16
                    else:
                        off = 16 * n + 10
                        buf = np.ascontiguousarray(buf.T)
                    a, b = buf.copy(), buf.copy()
                    L.or_lpf(a.ctypes.data, int(bd > 8), bd, off, n, vertical, taps, level, sharp)
                    mbl, lim, hev = oracle.lf_thresholds(level, sharp)
                    assert tab.lpf(_ptr(b), n * n, off, n, vertical, taps, mbl, lim, hev, bd) == 0
                    assert np.array_equal(a, b), (taps, vertical, level, sharp, trial)


def test_cdef_find_dir_and_filter_block(tab, oracle):
    rng = np.random.default_rng(60)
    L = oracle.lflib()
    for cs in (0, 2):
        mx = (1 << (8 + cs)) - 1
        for t in range(20):
            blk = rng.integers(0, mx + 1, (8, 8)).astype(np.uint16)
            if t % 3 == 0:
                blk = (np.add.outer(np.arange(8), np.arange(8) * (t % 5)) * 9 % (mx + 1)).astype(np.uint16)
            v = ctypes.c_int32()
            d = tab.cdef_find_dir(blk.ctypes.data_as(P(ctypes.c_uint16)), 8, cs, ctypes.byref(v))
            assert (d, v.value) == oracle.cdef_find_dir(blk, cs), (cs, t)
    BS = 144
    for bd in (8, 10):
        cs = bd - 8
        mx = (1 << bd) - 1
        for bsize, (bw, bh) in enumerate([(4, 4), (4, 8), (8, 4), (8, 8)]):
            for t in range(6):
                inb = rng.integers(0, mx + 1, (20, BS)).astype(np.uint16)
                if t % 2:
                    inb[:3, :] = 0x7f7f   # a frame edge above the block
                    inb[:, :9] = 0x7f7f
                ioff = 4 * BS + 16
                pri = int(rng.choice([0, 1, 5, 15])) << cs
                sec = int(rng.choice([0, 1, 2, 4])) << cs
                dirn = int(rng.integers(0, 8))
                pd, sd = 3 + cs + int(rng.integers(0, 3)), 4 + cs
                sub = 1 if t < 4 else 2
                ref = np.full((bh, 8), -1, np.int32)
                L.or_cdef_filter_block(ref.ctypes.data, 8, inb.ctypes.data + 2 * ioff, pri, sec, dirn, pd, sd, bw,
                                       bh, cs, sub)
                got = np.full((bh, 8), 0xffff if bd > 8 else 0xff, np.uint16 if bd > 8 else np.uint8)
                rc = tab.cdef_filter_block(_ptr(got), 8, inb.ctypes.data_as(P(ctypes.c_uint16)), inb.size, ioff, pri,
                                           sec, dirn, pd, sd, bsize, cs, sub, bd)
                assert rc == 0
                for i in range(0, bh, sub):
                    assert np.array_equal(got[i, :bw].astype(np.int32), ref[i, :bw]), (bd, bsize, t, i)
                for i in range(bh):
                    if i % sub:
                        assert np.all(got[i] == (0xffff if bd > 8 else 0xff))


@pytest.mark.parametrize('bd', [8, 10])
def test_compute_stats(tab, oracle, bd):
    rng = np.random.default_rng(70 + bd)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    pw, ph = 160, 120
    dgd = rng.integers(0, mx + 1, (ph, pw)).astype(dt)
    src = np.clip(dgd.astype(np.int64) + rng.integers(-9, 10, dgd.shape), 0, mx).astype(dt)
    for win in (7, 5):
        for (hs, he, vs, ve) in [(3, 67, 3, 67), (10, 150, 5, 100), (40, 41, 50, 51), (3, 157, 3, 117)]:
            M, H = oracle.wiener_stats(win, dgd, src, bd, (hs, he, vs, ve))
            m = np.zeros(win * win, np.int64)
            h = np.zeros((win * win) ** 2, np.int64)
            rc = tab.compute_stats(win, _ptr(dgd), dgd.size, 0, pw, _ptr(src), src.size, 0, pw, hs, he, vs, ve,
                                   m.ctypes.data_as(P(ctypes.c_int64)), h.ctypes.data_as(P(ctypes.c_int64)), bd)
            assert rc == 0
            assert np.array_equal(m, M) and np.array_equal(h.reshape(H.shape), H), (win, hs, he, vs, ve)
    # a window whose margin would start before the buffer is rejected
    m = np.zeros(49, np.int64)
    h = np.zeros(49 * 49, np.int64)
    assert tab.compute_stats(7, _ptr(dgd), dgd.size, 0, pw, _ptr(src), src.size, 0, pw, 0, 10, 1, 15,
                             m.ctypes.data_as(P(ctypes.c_int64)), h.ctypes.data_as(P(ctypes.c_int64)), bd) == -22


@pytest.mark.parametrize('kind', ['fp', 'b'])
def test_quantize_hbd_plain(tab, gpu, oracle, kind):
    """quantize_fp_hbd / quantize_b_hbd (ABI 3) against the oracle's 10-bit
    quantizers (no INT16 clamp) and the eob over the scan."""
    rng = np.random.default_rng(7 if kind == 'fp' else 8)
    f = tab.quantize_fp_hbd if kind == 'fp' else tab.quantize_b_hbd
    for tx_size, n, ls in ((0, 16, 0), (2, 256, 0), (3, 1024, 1), (4, 1024, 2), (9, 512, 1)):
        for sharp in (0, 3, -2):
            qi = int(rng.integers(1, 256))
            row = gpu.QuantRow()
            assert gpu.lib().zav1_build_quant_row_sharp(qi, 10, sharp, ctypes.byref(row)) == 0
            coef = (rng.normal(0, 1, n) * rng.choice([50, 800, 40000], n)).astype(np.int32)
            coef[rng.random(n) < 0.3] = 0
            scan = oracle.scan(tx_size, int(rng.integers(0, 3))).astype(np.uint16)
            q = np.zeros(n, np.int32)
            dq = np.zeros(n, np.int32)
            eob = f(coef.ctypes.data_as(P(ctypes.c_int32)), scan.ctypes.data_as(P(ctypes.c_uint16)), n,
                    ctypes.byref(row), ls, q.ctypes.data_as(P(ctypes.c_int32)), dq.ctypes.data_as(P(ctypes.c_int32)))
            orow = oracle.QuantRow()
            for fld, _ in oracle.QuantRow._fields_:
                setattr(orow, fld, type(getattr(orow, fld))(*getattr(row, fld)))
            rq, rdq = oracle.quantize(kind, coef, orow, ls, 10)
            nz = np.nonzero(rq[scan])[0]
            assert eob == (int(nz[-1]) + 1 if nz.size else 0), (tx_size, sharp)
            assert np.array_equal(q, rq) and np.array_equal(dq, rdq), (tx_size, sharp)


@pytest.mark.parametrize('bd', [8, 10])
def test_cdef_dist_plain(tab, oracle, bd):
    """cdef_dist (ABI 3) against the oracle's compute_cdef_dist."""
    L = oracle.lflib()
    L.or_cdef_dist.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_cdef_dist.restype = ctypes.c_uint64
    rng = np.random.default_rng(bd)
    dt = np.uint8 if bd == 8 else np.uint16
    mx = (1 << bd) - 1
    W = 72
    plane = rng.integers(0, mx + 1, (64, W)).astype(dt)
    for bsize, (bw, bh) in enumerate(((4, 4), (4, 8), (8, 4), (8, 8))):
        nb = int(rng.integers(1, 17))
        dl = np.stack([rng.integers(0, 48 // bh, nb), rng.integers(0, 64 // bw, nb)], 1).astype(np.uint8)
        packed = np.clip(rng.integers(-20, 21, (nb, bh, bw)) + mx // 2, 0, mx).astype(dt)
        for sub in (1, 2):
            cs = bd - 8
            out = ctypes.c_uint64()
            off = 3 * W + 5
            rc = tab.cdef_dist(_ptr(plane), plane.size, off, W, _ptr(packed), packed.size, _ptr(dl), nb, bsize, cs,
                               sub, bd, ctypes.byref(out))
            assert rc == 0
            ref = L.or_cdef_dist(plane.ctypes.data, int(bd > 8), off, W, packed.ctypes.data, dl.ctypes.data, nb,
                                 bsize, cs, sub)
            assert out.value == ref, (bd, bsize, sub)
