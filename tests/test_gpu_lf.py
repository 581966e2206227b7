"""GPU parity of the loop-filter stage against the oracle, bit-exact:
deblocking (all planes, partial frames, skip maps, levels, sharpness),
the CDEF direction + 64-strength SSE tables per filter block (luma and joint
chroma, row subsampling, partial blocks), Wiener M/H, unit SSE and
self-guided projection statistics of every restoration unit; and the
whole stage at the 4096x4096 8-bit workload size with sampled checks."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod(name):
    path = os.path.join(ROOT, 'zenav1-svt_amd', name + '.py')
    spec = importlib.util.spec_from_file_location('zav1_' + name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


synth = _mod('synth')


def blocky(h, w, bd, seed, blk=8):
    """Noise with flattened blocks and block-edge steps (deblock/CDEF food)."""
    rng = np.random.default_rng(seed)
    f = synth.spectral_noise(h, w, 1.0, seed)
    mx = (1 << bd) - 1
    src = np.clip(np.round(f * mx * 0.8 + mx * 0.1), 0, mx).astype(np.int32)
    rec = src.copy()
    for y in range(0, h - blk + 1, blk):
        for x in range(0, w - blk + 1, blk):
            if rng.uniform() < 0.4:
                rec[y:y + blk, x:x + blk] = int(rec[y:y + blk, x:x + blk].mean())
    rec = rec + rng.integers(-3 << (bd - 8), (3 << (bd - 8)) + 1, rec.shape)
    dt = np.uint8 if bd == 8 else np.uint16
    return src.astype(dt), np.clip(rec, 0, mx).astype(dt)


@pytest.mark.parametrize('bd,skip,sharp', [(8, 0.0, 0), (10, 0.0, 0), (8, 0.4, 3), (10, 0.4, 5)])
def test_deblock_planes_bitexact(gpu, oracle, bd, skip, sharp):
    z = gpu
    H, W = 136, 200
    mi_rows, mi_cols = H // 4, W // 4
    true_w, true_h = W - 2, H - 2
    geom = synth.partition_geometry(mi_rows, mi_cols, 4000 + bd, skip_frac=skip)
    gd = {k: z.DeviceArray.from_host(v) for k, v in geom.items()}
    g = z.deblock_geom(mi_cols, mi_rows, true_w, true_h, gd)
    for plane, (ph, pw, ss) in enumerate([(H, W, 0), (H // 2, W // 2, 1), (H // 2, W // 2, 1)]):
        _, rec = blocky(ph, pw, bd, 50 + plane, blk=8 if plane == 0 else 4)
        for lv, lh in [(40, 24), (63, 63), (0, 17), (9, 0)]:
            d = z.DeviceArray.from_host(rec)
            z.deblock_plane(d, pw, pw, ph, bd, plane, ss, lv, lh, sharp, g)
            got = d.to_host()
            ref = oracle.deblock_plane(rec, bd, plane, ss, lv, lh, sharp, geom, mi_cols, mi_rows, true_w, true_h)
            assert not np.array_equal(ref, rec) or (lv == 0 and lh == 0)
            if not np.array_equal(got, ref):
                bad = np.argwhere(got != ref)[:4]
                pytest.fail('plane %d levels %d/%d: %d mismatches, first %s' % (
                    plane, lv, lh, int((got != ref).sum()), bad.tolist()))


@pytest.mark.parametrize('bd,sub,qindex,skipf', [(8, 1, 100, 0.0), (10, 1, 200, 0.0), (8, 2, 40, 0.5), (10, 4, 255, 0.6)])
def test_cdef_search_bitexact(gpu, oracle, bd, sub, qindex, skipf):
    z = gpu
    H, W = 136, 200  # 4 x 3 filter blocks, the last row / column partial
    skip = synth.partition_geometry(H // 4, W // 4, 77, skip_frac=skipf)['skip'] if skipf else None
    sy, ry = blocky(H, W, bd, 70, 8)
    su, ru = blocky(H // 2, W // 2, bd, 71, 4)
    sv, rv = blocky(H // 2, W // 2, bd, 72, 4)
    D = z.DeviceArray
    rec = [D.from_host(a) for a in (ry, ru, rv)]
    src = [D.from_host(a) for a in (sy, su, sv)]
    nhfb, nvfb = (W + 63) // 64, (H + 63) // 64
    dv = D((nhfb * nvfb, 64, 2), np.int32)
    my = D((nhfb * nvfb, 64), np.uint64)
    muv = D((nhfb * nvfb, 64), np.uint64)
    dsk = D.from_host(skip) if skip is not None else None
    z.cdef_search(rec, src, W, H, bd, qindex, sub, dv, my, muv, d_skip=dsk, mi_cols=W // 4, mi_rows=H // 4)
    gy, guv, gdv = my.to_host(), muv.to_host(), dv.to_host()
    n_skipped = 0
    fs = np.arange(64)
    for fbr in range(nvfb):
        for fbc in range(nhfb):
            fb = fbr * nhfb + fbc
            ey, euv = oracle.cdef_search_fb((ry, ru, rv), (sy, su, sv), W, H, bd, qindex, sub, fbr, fbc, fs, skip=skip)
            assert np.array_equal(gy[fb], ey), (fb, np.nonzero(gy[fb] != ey)[0][:8], gy[fb][:4], ey[:4])
            assert np.array_equal(guv[fb], euv), (fb, np.nonzero(guv[fb] != euv)[0][:8])
            vs, hs = min(64, H - fbr * 64), min(64, W - fbc * 64)
            for by in range(vs // 8):
                for bx in range(hs // 8):
                    mr, mc = fbr * 16 + by * 2, fbc * 16 + bx * 2
                    if skip is not None and skip[mr:mr + 2, mc:mc + 2].all():
                        n_skipped += 1
                        assert tuple(gdv[fb, by * 8 + bx]) == (0, 0)
                        continue
                    blk = ry[fbr * 64 + by * 8:fbr * 64 + by * 8 + 8, fbc * 64 + bx * 8:fbc * 64 + bx * 8 + 8]
                    d, v = oracle.cdef_find_dir(blk, bd - 8)
                    assert tuple(gdv[fb, by * 8 + bx]) == (d, v), (fb, by, bx)
    assert (n_skipped > 0) == (skip is not None)
    # the table is not degenerate: some strength beats no filtering
    assert (gy.min(axis=1) < gy[:, 0]).any()


def _units(oracle, us, pw, ph, ss):
    nh = oracle.lrlib().or_count_units(us, pw)
    nv = oracle.lrlib().or_count_units(us, ph)
    return nh, nv, [oracle.unit_limits(us, pw, ph, ss, ux, uy) for uy in range(nv) for ux in range(nh)]


@pytest.mark.parametrize('bd', [8, 10])
@pytest.mark.parametrize('win,us,ss,dims', [(7, 64, 0, (200, 300)), (5, 32, 1, (100, 150)), (7, 128, 0, (136, 200))])
def test_wiener_stats_and_unit_sse_bitexact(gpu, oracle, bd, win, us, ss, dims):
    z = gpu
    ph, pw = dims
    src, rec = blocky(ph, pw, bd, 90 + win, 8)
    nh, nv, lims = _units(oracle, us, pw, ph, ss)
    assert z.rest_unit_grid(us, pw, ph) == (nh, nv)
    D = z.DeviceArray
    dd, ds = D.from_host(rec), D.from_host(src)
    w2 = win * win
    M, Hm, sse = D((nh * nv, w2), np.int64), D((nh * nv, w2, w2), np.int64), D((nh * nv,), np.uint64)
    z.wiener_stats(dd, ds, pw, ph, bd, win, us, ss, M, Hm)
    z.unit_sse(dd, ds, pw, ph, bd, us, ss, sse)
    gm, gh, gs = M.to_host(), Hm.to_host(), sse.to_host()
    L = oracle.lrlib()
    for u, lim in enumerate(lims):
        em, eh = oracle.wiener_stats(win, rec, src, bd, lim)
        assert np.array_equal(gm[u], em), (u, lim)
        assert np.array_equal(gh[u], eh), (u, lim, np.argwhere(gh[u] != eh)[:4].tolist())
        es = L.or_sse_region(rec.ctypes.data, src.ctypes.data, pw, bd, *lim)
        assert int(gs[u]) == es, (u, lim)


@pytest.mark.parametrize('bd', [8, 10])
def test_sgr_stats_bitexact(gpu, oracle, bd):
    z = gpu
    ph, pw, us = 100, 160, 64
    src, rec = blocky(ph, pw, bd, 120 + bd, 8)
    nh, nv, lims = _units(oracle, us, pw, ph, 0)
    D = z.DeviceArray
    out = D((nh * nv, 16, 5), np.int64)
    z.sgr_stats(D.from_host(rec), D.from_host(src), pw, ph, bd, us, 0, out)
    got = out.to_host()
    for u, lim in enumerate(lims):
        for s in range(16):
            ref = oracle.sgr_unit_stats(rec, src, bd, lim, s)
            assert np.array_equal(got[u, s], ref), (u, s, got[u, s], ref)


def _sgr_content(kind, h, w, bd, seed):
    rng = np.random.default_rng(seed)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    if kind == 'extreme':
        # 0 / max in 2x2 cells: box variances at the largest the bit depth
        # allows (the u32 products of the scale step at their bound)
        cells = rng.integers(0, 2, ((h + 1) // 2, (w + 1) // 2)) * mx
        rec = np.kron(cells, np.ones((2, 2), np.int64))[:h, :w]
        src = rng.integers(0, mx + 1, (h, w))
    elif kind == 'noise':
        rec, src = rng.integers(0, mx + 1, (h, w)), rng.integers(0, mx + 1, (h, w))
    else:  # flat halves next to noise: z = 0 and saturated z side by side
        rec = np.full((h, w), mx // 3)
        rec[:, w // 2:] = rng.integers(0, mx + 1, (h, w - w // 2))
        src = np.clip(rec + rng.integers(-9, 10, (h, w)), 0, mx)
    return src.astype(dt), rec.astype(dt)


@pytest.mark.parametrize('pw,ph,us,ss,bd,kind', [
    (200, 70, 32, 1, 10, 'extreme'),   # strips spanning three units, chroma offsets
    (200, 70, 32, 1, 8, 'extreme'),
    (130, 9, 64, 0, 8, 'noise'),       # one short unit row
    (300, 140, 128, 0, 10, 'flat'),
    (59, 61, 256, 0, 10, 'noise'),     # one unit narrower than a strip
    (117, 200, 64, 0, 8, 'flat'),      # last unit row extended, odd height
])
def test_sgr_stats_shapes_and_extremes(gpu, oracle, pw, ph, us, ss, bd, kind):
    """All 16 parameter sets of every unit equal the oracle across unit sizes
    smaller and larger than a 58-column strip, ragged planes, chroma unit
    offsets and contents at the arithmetic bounds."""
    z = gpu
    src, rec = _sgr_content(kind, ph, pw, bd, pw * 7 + ph)
    nh, nv, lims = _units(oracle, us, pw, ph, ss)
    D = z.DeviceArray
    out = D((nh * nv, 16, 5), np.int64)
    z.sgr_stats(D.from_host(rec), D.from_host(src), pw, ph, bd, us, ss, out)
    got = out.to_host()
    for u, lim in enumerate(lims):
        for s in range(16):
            ref = oracle.sgr_unit_stats(rec, src, bd, lim, s)
            assert np.array_equal(got[u, s], ref), (u, lim, s, got[u, s], ref)


def test_loop_filter_stage_full_size_sampled(gpu, oracle):
    """The 4096x4096 8-bit workload end to end; deblocked planes checked in
    full, the CDEF / restoration tables on sampled blocks and units."""
    z = gpu
    lfm = _mod('loopfilter')
    src_y, rec_y = synth.config3_recon(4096, 4096, seed=3000)
    (su, ru), (sv, rv) = synth.config3_chroma(2048, 2048, seed=3000)
    geom = synth.partition_geometry(1024, 1024, 3001, skip_frac=0.2)
    deb = {k: v for k, v in geom.items() if k != 'skip'}
    qindex = synth.qp_to_qindex(32)
    levels = synth.key_frame_lf_levels(int(z.quant_row(qindex, 8).dequant[1]), 8)
    st = lfm.LoopFilterStage(z, {'src': (src_y, su, sv), 'rec': (rec_y, ru, rv)}, geom, 8, qindex, levels)
    st.run()
    z.sync()
    dy = [d.to_host() for d in st.dgd]
    ref = [oracle.deblock_plane(rec_y, 8, 0, 0, levels[0], levels[1], 0, deb, 1024, 1024, 4096, 4096),
           oracle.deblock_plane(ru, 8, 1, 1, levels[2], levels[2], 0, deb, 1024, 1024, 4096, 4096),
           oracle.deblock_plane(rv, 8, 2, 1, levels[3], levels[3], 0, deb, 1024, 1024, 4096, 4096)]
    for p in range(3):
        assert np.array_equal(dy[p], ref[p]), p
    gy, guv = st.mse_y.to_host(), st.mse_uv.to_host()
    rng = np.random.default_rng(7)
    for fb in list(rng.integers(0, 4096, 6)) + [0, 4095]:
        fbr, fbc = divmod(int(fb), 64)
        ey, euv = oracle.cdef_search_fb(tuple(ref), (src_y, su, sv), 4096, 4096, 8, qindex, 1, fbr, fbc,
                                        np.arange(64), skip=geom['skip'])
        assert np.array_equal(gy[fb], ey) and np.array_equal(guv[fb], euv), fb
    for p, (pw, ss, srcp) in enumerate([(4096, 0, src_y), (2048, 1, su), (2048, 1, sv)]):
        nh, nv = st.grid[p]
        gm, gh, gsgr = st.M[p].to_host(), st.H[p].to_host(), st.sgrs[p].to_host()
        for u in [0, nh * nv - 1, int(rng.integers(0, nh * nv))]:
            lim = oracle.unit_limits(st.ru[p], pw, pw, ss, u % nh, u // nh)
            em, eh = oracle.wiener_stats(st.win[p], dy[p], srcp, 8, lim)
            assert np.array_equal(gm[u], em) and np.array_equal(gh[u], eh), (p, u)
            for s in (0, 10, 14):
                assert np.array_equal(gsgr[u, s], oracle.sgr_unit_stats(dy[p], srcp, 8, lim, s)), (p, u, s)


@pytest.mark.parametrize('bd', [8, 10])
def test_deblock_level_trials_and_search(gpu, oracle, bd):
    """Trial SSE of every candidate level and the hill-climb pick equal the
    oracle's sequential search with real trials, luma and chroma, both
    convergence settings."""
    z = gpu
    H, W = 136, 200
    mi_rows, mi_cols = H // 4, W // 4
    geom = synth.partition_geometry(mi_rows, mi_cols, 600 + bd)
    gd = {k: z.DeviceArray.from_host(v) for k, v in geom.items()}
    g = z.deblock_geom(mi_cols, mi_rows, W, H, gd)
    for plane, (ph, pw, ss) in enumerate([(H, W, 0), (H // 2, W // 2, 1)]):
        # block means of a gentle ramp: small steps the filters can smooth
        yy, xx = np.mgrid[0:ph, 0:pw]
        dt = np.uint8 if bd == 8 else np.uint16
        src = ((100 + 0.5 * xx + 0.25 * yy + 2 * np.sin(xx / 5.0)) * (1 << (bd - 8))).astype(dt)
        blk = 8 if plane == 0 else 4
        rec = src.astype(np.int64).copy()
        for y in range(0, ph, blk):
            for x in range(0, pw, blk):
                rec[y:y + blk, x:x + blk] = rec[y:y + blk, x:x + blk].mean()
        rec = rec.astype(dt)
        d_rec, d_src = z.DeviceArray.from_host(rec), z.DeviceArray.from_host(src)
        levels = [0, 3, 9, 17, 32, 47, 63]
        d_work = z.DeviceArray((len(levels), ph, pw), src.dtype)
        d_sse = z.DeviceArray(64, np.uint64)
        z.deblock_level_sse(d_rec, d_src, pw, ph, bd, plane, ss, 2, g, levels, d_work, d_sse)
        got = d_sse.to_host()[:len(levels)]
        for i, lv in enumerate(levels):
            f = oracle.deblock_plane(rec, bd, plane, ss, lv, lv, 2, geom, mi_cols, mi_rows, W, H) if lv else rec
            ref = int(((f.astype(np.int64) - src.astype(np.int64)) ** 2).sum())
            assert int(got[i]) == ref, (plane, lv, int(got[i]), ref)
        d_work2 = z.DeviceArray((2, ph, pw), src.dtype)
        for conv in (0, 1):
            res = z.deblock_search_level(d_rec, d_src, pw, ph, bd, plane, ss, 2, g, conv, d_work2, d_sse)
            ref = oracle.deblock_search_level(rec, src, bd, plane, ss, 2, geom, mi_cols, mi_rows, W, H, conv)
            assert res == ref, (plane, conv, res, ref)
            assert res[0] > 0
