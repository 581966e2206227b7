"""GPU checks of the wavefront serving loop (continuous batching with HBM
admission): requests of mixed size, bit depth, qp and preset arrive faster
than they complete, overlap on the device, are held back when the budget is
full, and every request's reconstruction and decisions in the mix equal
the same request run alone, and the oracle's sequential restatement."""
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


FIELDS = ('cand', 'tx_type', 'eob', 'rate', 'mode_rate', 'dist', 'cost')
REQS = [(128, 192, 10, 20, 4), (64, 128, 8, 45, 6), (192, 64, 10, 33, 8), (256, 256, 8, 28, 4), (64, 512, 10, 50, 6),
        (128, 128, 8, 38, 8), (320, 192, 10, 24, 6), (64, 64, 8, 30, 4), (256, 128, 10, 42, 4), (128, 320, 8, 22, 8)]


def test_serving_mix_alone_equals_in_mix(gpu, oracle):
    z = gpu
    synth, wf, srv = _mod('synth'), _mod('wavefront'), _mod('serving')
    imgs = synth.serving_images(REQS, seed=77)
    specs = [(img, bd, qp, p) for img, (h, w, bd, qp, p) in zip(imgs, REQS)]
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(0.002, len(REQS)))
    probe = srv.WavefrontServer(z, wf, synth, 1 << 40)
    fps = sorted(wf.WavefrontMD.footprint(z, probe.request_cfg(h, w, bd, qp, p)) for h, w, bd, qp, p in REQS)
    budget = fps[-1] + fps[-2]  # the two largest fit together, not three
    server = srv.WavefrontServer(z, wf, synth, budget, max_active=4, chunk=8, depth=2)
    done, st = server.run(specs, arrivals)
    server.close()
    assert st['peak_active'] >= 2, st
    assert st['deferred_by_memory'] >= 1, st
    assert st['peak_bytes'] <= budget
    lat = st['latency_s']
    assert lat.shape == (len(REQS),) and np.all(lat > 0)
    for r in done:
        md = wf.WavefrontMD(z, r.cfg)
        try:
            md.upload(r.img[None])
            md.run(use_graph=False)
            z.sync()
            assert np.array_equal(md.recon()[0], r.recon), r.idx
            alone = md.decisions()[0]
            for f in FIELDS:
                assert np.array_equal(alone[f], r.decisions[f]), (r.idx, f)
        finally:
            md.close()
    # the oracle for the 8-bit and 10-bit requests at each preset
    for i in (0, 1, 2, 5, 7):
        r = done[i]
        orec, odec = oracle.wf_frame(r.cfg, r.img[None].astype(np.uint8 if r.bd == 8 else np.uint16))
        assert np.array_equal(orec[0], r.recon), i
        for f in FIELDS:
            assert np.array_equal(odec[0][f], r.decisions[f]), (i, f)


def test_serving_rejects_a_request_over_budget(gpu):
    z = gpu
    synth, wf, srv = _mod('synth'), _mod('wavefront'), _mod('serving')
    img = synth.serving_images([(256, 256, 10, 30, 4)], seed=3)[0]
    server = srv.WavefrontServer(z, wf, synth, 1 << 16)  # below the plane's own 128 KiB
    with pytest.raises(ValueError):
        server.run([(img, 10, 30, 4)], [0.0])


def test_serving_encode_streams_alone_equal_in_mix(gpu, oracle):
    """Encode mode: every request's AV1 stream, written on the host pool
    while other requests run, equals the stream of the same request decided
    and written alone, and decodes (the oracle's decoder) to its
    reconstruction; a 4096x3072 request takes two tile columns (AV1's tile
    area limit)."""
    z = gpu
    synth, wf, srv = _mod('synth'), _mod('wavefront'), _mod('serving')
    reqs = REQS[:6] + [(3072, 4096, 10, 36, 8)]
    imgs = synth.serving_images(reqs, seed=78)
    specs = [(img, bd, qp, p) for img, (h, w, bd, qp, p) in zip(imgs, reqs)]
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(0.002, len(reqs)))
    server = srv.WavefrontServer(z, wf, synth, 8 << 30, max_active=4, chunk=16, depth=2, encode=True)
    done, st = server.run(specs, arrivals)
    server.close()
    assert st['stream_bytes'] == sum(len(r.stream_bytes) for r in done) > 0
    assert int(done[-1].cfg[0]['tile_cols']) == 2
    for r in done:
        md = wf.WavefrontMD(z, r.cfg)
        try:
            md.upload(r.img[None])
            md.run(use_graph=False)
            z.sync()
            alone = md.write_stills(0, 1, threads=2)[0]
        finally:
            md.close()
        assert alone == r.stream_bytes, r.idx
        h, w = r.img.shape
        frame, info, _, _ = oracle.av1_decode_still(r.stream_bytes, w, h, r.bd)
        assert np.array_equal(frame, r.recon), r.idx
    print('%d requests, %d stream bytes' % (len(done), st['stream_bytes']), flush=True)
