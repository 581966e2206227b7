"""GPU checks of the serving loop: every request completes with a latency,
and a request's arena results equal the oracle's search at its qp."""
import ctypes
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
    sys.modules['zav1_' + name] = m
    spec.loader.exec_module(m)
    return m


def test_serving_loop_latencies_and_results(gpu, oracle):
    zav1 = gpu
    synth, tsm, srv = _mod('synth'), _mod('txsearch'), _mod('serving')
    reqs = [(128, 192, 10, 20), (64, 128, 8, 45), (192, 64, 10, 33), (256, 256, 8, 28), (64, 4096, 10, 50)]
    imgs = synth.serving_images(reqs, seed=77)
    loop = srv.ServingLoop(zav1, tsm, synth, n_ctx=2, ctx_budget=srv.request_bytes(256, 4096, tsm))
    lat, mp, wall = loop.run(reqs, imgs, rate_hz=2000.0)
    assert lat.shape == (5,) and np.all(lat > 0) and wall > 0
    assert abs(mp - sum(h * w for h, w, _, _ in reqs) / 1e6) < 1e-9
    # one request through a context, results read back from the arena
    h, w, bd, qp = reqs[1]
    fs = loop.submit(0, imgs[1], bd, synth.qp_to_qindex(qp))
    zav1.sync(loop.ctx[0].stream())
    L = zav1.lib()
    for ts in range(19):
        nb = fs.n_blocks[ts]
        types = np.array(tsm.legal_types(ts), np.uint8)
        assert fs.n_types[ts] == types.size and list(fs.types[ts])[:types.size] == list(types)
        assert nb == (w // tsm.TX_DIMS[ts][0]) * (h // tsm.TX_DIMS[ts][1])
        if nb == 0:
            assert not fs.dist[ts]
            continue
        dd, de, dl = fs.dist[ts], fs.eob[ts], fs.lsum[ts]
        n = nb * types.size
        d, e, l = np.zeros(n, np.uint64), np.zeros(n, np.int32), np.zeros(n, np.uint32)
        for host, dev in ((d, dd), (e, de), (l, dl)):
            zav1.check(L.zav1_memcpy_d2h(host.ctypes.data, dev, host.nbytes, None), 'd2h')
        xy = tsm.tile_xy(w, h, ts)
        rd, re, rl = oracle.txsearch_plane(imgs[1], bd, ts, xy, types, synth.qp_to_qindex(qp))
        assert np.array_equal(d, rd) and np.array_equal(e, re) and np.array_equal(l, rl), ts
    # an arena too small for a request: ENOMEM before anything is enqueued
    small = zav1.Context(0, 1 << 16)
    d_plane = small.upload(imgs[3][:64, :64].copy(), 10)
    q = np.array([80], np.int32)
    f2 = zav1.FrameSearch()
    assert L.zav1_txsearch_frame(small.h, d_plane, 256, 256, 8, q.ctypes.data, 1, ctypes.byref(f2)) == zav1.ZAV1_ENOMEM
    assert small.usage()[0] == 64 * 64 * 2
    small.close()
    loop.close()
