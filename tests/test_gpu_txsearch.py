"""GPU parity of the fused tx-type search (DC prediction -> residual -> fwd ->
fold -> fp quantize -> distortion / eob / level sum) against the oracle."""
import importlib.util
import os
import sys

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


@pytest.mark.parametrize('bd,qp', [(10, 20), (10, 40), (8, 32)])
def test_txsearch_all_sizes_bitexact(gpu, oracle, bd, qp):
    zav1 = gpu
    synth, ts = _mod('synth'), _mod('txsearch')
    plane = synth.config2_luma(256, 256, seed=2000 + qp)
    if bd == 8:
        plane = (plane >> 2).astype(np.uint16)
    qi = synth.qp_to_qindex(qp)
    sweep = ts.TxSearchSweep(zav1, plane, bd, [qi])
    sweep.run()
    zav1.sync()
    for tx_size in range(19):
        xy, types, dist, eob, lsum = sweep.results(tx_size)
        rd, re, rl = oracle.txsearch_plane(plane, bd, tx_size, xy, types, qi)
        assert np.array_equal(dist, rd), (tx_size, np.argwhere(dist != rd)[:3])
        assert np.array_equal(eob, re), tx_size
        assert np.array_equal(lsum, rl), tx_size


def test_txsearch_ragged_and_edges(gpu, oracle):
    """Blocks on the frame's first row/column (no neighbours), a ragged pair
    count, extreme samples."""
    zav1 = gpu
    rng = np.random.default_rng(9)
    plane = rng.integers(0, 1024, (96, 160)).astype(np.uint16)
    plane[:8, :8] = 1023
    plane[8:16, :8] = 0
    xy = np.array([[0, 0], [0, 8], [8, 0], [24, 40], [152, 88], [100, 60], [4, 4]], np.int32)
    for tx_size in (0, 1, 5, 6, 13):
        w, h = zav1.TX_DIMS[tx_size]
        ok = (xy[:, 0] + w <= 160) & (xy[:, 1] + h <= 96)
        sub = xy[ok]
        types = np.array([0, 3, 5, 9, 15], np.uint8)
        d_plane = zav1.DeviceArray.from_host(plane)
        d_xy = zav1.DeviceArray.from_host(sub)
        n = sub.shape[0] * types.size
        dd, de, dl = zav1.DeviceArray(n, np.uint64), zav1.DeviceArray(n, np.int32), zav1.DeviceArray(n, np.uint32)
        zav1.txsearch_batch(d_plane, 160, 10, tx_size, d_xy, sub.shape[0], types, 255, dd, de, dl)
        rd, re, rl = oracle.txsearch_plane(plane, 10, tx_size, sub, types, 255)
        assert np.array_equal(dd.to_host(), rd) and np.array_equal(de.to_host(), re)
        assert np.array_equal(dl.to_host(), rl)


@pytest.mark.parametrize('bd', [8, 10])
def test_txsearch_type_orders_and_extremes(gpu, oracle, bd):
    """Caller type lists in any order (one type, all 16 reversed, a subset
    without some column variants) on extreme content: checkerboards and
    full-range steps at qindex 0 (largest levels and errors) and 255."""
    zav1 = gpu
    rng = np.random.default_rng(bd)
    top = (1 << bd) - 1
    plane = rng.integers(0, top + 1, (64, 128)).astype(np.uint16)
    yy, xx = np.mgrid[0:64, 0:128]
    plane[:32, :64] = np.where((yy[:32, :64] + xx[:32, :64]) % 2 == 0, top, 0)
    plane[32:, :64] = np.where(xx[32:, :64] % 8 < 4, top, 0)
    d_plane = zav1.DeviceArray.from_host(plane)
    lists = [[9], list(range(15, -1, -1)), [4, 6, 8, 14, 1], [10, 11, 12, 13, 14, 15]]
    for tx_size in range(19):
        w, h = zav1.TX_DIMS[tx_size]
        ys, xs = np.mgrid[0:64 - h + 1:h, 0:128 - w + 1:w]
        xy = np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.int32)
        d_xy = zav1.DeviceArray.from_host(xy)
        for tl in lists:
            if max(w, h) == 64:
                tl = [0]
            elif max(w, h) == 32:
                tl = [t for t in tl if t in (0, 9, 10, 11)] or [10, 0]
            types = np.array(tl, np.uint8)
            for qi in (0, 255):
                n = xy.shape[0] * types.size
                dd, de, dl = (zav1.DeviceArray(n, np.uint64), zav1.DeviceArray(n, np.int32),
                              zav1.DeviceArray(n, np.uint32))
                zav1.txsearch_batch(d_plane, 128, bd, tx_size, d_xy, xy.shape[0], types, qi, dd, de, dl)
                rd, re, rl = oracle.txsearch_plane(plane, bd, tx_size, xy, types, qi)
                assert np.array_equal(dd.to_host(), rd), (tx_size, tl, qi)
                assert np.array_equal(de.to_host(), re), (tx_size, tl, qi)
                assert np.array_equal(dl.to_host(), rl), (tx_size, tl, qi)


@pytest.mark.parametrize('bd', [8, 10])
def test_txsearch_multi_quantizer_matches_single(gpu, oracle, bd):
    """The several-quantizer entry point (transforms shared, one launch per
    pair of quantizers) equals one call per qindex and the oracle, including
    qindex 0 (the 32-bit level form at 10 bits) mixed with others and an odd
    quantizer count."""
    zav1 = gpu
    synth, ts = _mod('synth'), _mod('txsearch')
    plane = synth.config2_luma(128, 128, seed=77 + bd)
    if bd == 8:
        plane = (plane >> 2).astype(np.uint16)
    qis = [0, 60, 255]
    d_plane = zav1.DeviceArray.from_host(plane)
    for tx_size in range(19):
        xy = ts.tile_xy(128, 128, tx_size)
        types = np.array(ts.legal_types(tx_size), np.uint8)
        d_xy = zav1.DeviceArray.from_host(xy)
        n = xy.shape[0] * types.size
        dd, de, dl = (zav1.DeviceArray(3 * n, np.uint64), zav1.DeviceArray(3 * n, np.int32),
                      zav1.DeviceArray(3 * n, np.uint32))
        zav1.txsearch_batch_q(d_plane, 128, bd, tx_size, d_xy, xy.shape[0], types, qis, dd, de, dl)
        md, me, ml = dd.to_host(), de.to_host(), dl.to_host()
        for qp, qi in enumerate(qis):
            sd, se, sl = zav1.DeviceArray(n, np.uint64), zav1.DeviceArray(n, np.int32), zav1.DeviceArray(n, np.uint32)
            zav1.txsearch_batch(d_plane, 128, bd, tx_size, d_xy, xy.shape[0], types, qi, sd, se, sl)
            rd, re, rl = oracle.txsearch_plane(plane, bd, tx_size, xy, types, qi)
            sl_ = slice(qp * n, (qp + 1) * n)
            assert np.array_equal(md[sl_], rd) and np.array_equal(sd.to_host(), rd), (tx_size, qi)
            assert np.array_equal(me[sl_], re) and np.array_equal(se.to_host(), re), (tx_size, qi)
            assert np.array_equal(ml[sl_], rl) and np.array_equal(sl.to_host(), rl), (tx_size, qi)


def test_txsearch_frame_in_an_arena_and_enomem(gpu, oracle):
    """zav1_txsearch_frame: grids and outputs carved from a context's arena
    equal the oracle's search; an arena too small for a request gives
    ZAV1_ENOMEM with nothing enqueued (the arena keeps only the plane)."""
    import ctypes
    zav1 = gpu
    synth, tsm = _mod('synth'), _mod('txsearch')
    img = synth.serving_images([(64, 128, 8, 45, 6)], seed=77)[0]
    h, w, bd, qidx = 64, 128, 8, synth.qp_to_qindex(45)
    ctx = zav1.Context(0, 64 << 20)
    d_plane = ctx.upload(img, 10)  # host samples as u16 (the request's 8-bit content in 16-bit words)
    q = np.array([qidx], np.int32)
    fs = zav1.FrameSearch()
    L = zav1.lib()
    zav1.check(L.zav1_txsearch_frame(ctx.h, d_plane, w, h, bd, q.ctypes.data, 1, ctypes.byref(fs)), 'txsearch_frame')
    zav1.sync(ctx.stream())
    for ts in range(19):
        nb = fs.n_blocks[ts]
        types = np.array(tsm.legal_types(ts), np.uint8)
        assert fs.n_types[ts] == types.size and nb == (w // tsm.TX_DIMS[ts][0]) * (h // tsm.TX_DIMS[ts][1])
        if nb == 0:
            continue
        n = nb * types.size
        d, e, l = np.zeros(n, np.uint64), np.zeros(n, np.int32), np.zeros(n, np.uint32)
        for host, dev in ((d, fs.dist[ts]), (e, fs.eob[ts]), (l, fs.lsum[ts])):
            zav1.check(L.zav1_memcpy_d2h(host.ctypes.data, dev, host.nbytes, None), 'd2h')
        rd, re, rl = oracle.txsearch_plane(img, bd, ts, tsm.tile_xy(w, h, ts), types, qidx)
        assert np.array_equal(d, rd) and np.array_equal(e, re) and np.array_equal(l, rl), ts
    ctx.close()
    small = zav1.Context(0, 1 << 16)
    d_plane = small.upload(img[:64, :64].copy(), 10)
    f2 = zav1.FrameSearch()
    assert L.zav1_txsearch_frame(small.h, d_plane, 256, 256, 8, q.ctypes.data, 1, ctypes.byref(f2)) == zav1.ZAV1_ENOMEM
    assert small.usage()[0] == 64 * 64 * 2
    small.close()
