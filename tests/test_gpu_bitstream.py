"""End to end on the GPU: the wavefront mode decision (keep_levels) ->
zav1_wf_write_stills (host AV1 writer) -> the oracle's decoder written from
the specification.  The decoded frame must equal the device reconstruction
sample for sample and the decoded syntax the device decisions (mode, filter
intra, tx type, eob per leaf): the stream is a valid AV1 monochrome still of
exactly what the mode decision chose.  Uniform 16x16 leaves at 10 bits
(preset 4 funnel, 2 images), the PD0 trees at preset 8 (config 1), SB128
with 2x2 tiles at preset 6."""
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


wf = _mod('wavefront')
part = _mod('partition')
synth = _mod('synth')


def check_streams(z, O, md, images, bd):
    st = z.stream_create()
    md.run(False, st)
    z.stream_sync(st)
    rec = md.recon()
    streams = md.write_stills(0, len(images), threads=4, stream=st)
    dec = md.decisions_cells()
    z.stream_destroy(st)
    H, W = images[0].shape
    total = 0
    for i, s in enumerate(streams):
        frame, info, blocks, levels = O.av1_decode_still(s, W, H, bd)
        assert np.array_equal(frame, rec[i]), 'image %d: decoded frame != device reconstruction' % i
        for b in blocks:
            d = dec[i, b['y'] // 8, b['x'] // 8]
            assert (b['mode'], b['fi_mode'], b['eob']) == (d['mode'], d['fi_mode'], d['eob'])
            if b['eob']:
                assert b['tx_type'] == d['tx_type']
        total += len(s)
    print('%d images, %d bytes' % (len(images), total), flush=True)
    return streams


def test_uniform_leaves_10bit(gpu, oracle):
    imgs = [synth.config2_luma(256, 256, seed=2200 + k) for k in range(2)]
    qidx = synth.qp_to_qindex(40)
    cfg = wf.funnel_cfg(256, 256, 10, qidx, int(gpu.lib().zav1_kf_full_lambda(qidx, 40, 10)), 4, 40, n_images=2)
    cfg[0]['keep_levels'] = 1
    md = wf.WavefrontMD(gpu, cfg)
    md.upload(np.stack(imgs))
    check_streams(gpu, oracle, md, imgs, 10)
    md.close()


def test_pd0_tree_config1(gpu, oracle):
    img = synth.config1_luma(512, 512)
    qidx = synth.qp_to_qindex(32)
    pc = part.pd0_config(gpu, 512, 512, 8, 32, qidx, 8)
    d = gpu.DeviceArray.from_host(img)
    pp = part.PartitionPass(gpu, pc)
    pp.run(d)
    gpu.sync()
    lm = np.ascontiguousarray(pp.results()[2].reshape(-1), np.uint8)
    pp.close()
    d.free()
    cfg = wf.tree_cfg(512, 512, 8, qidx, int(gpu.lib().zav1_kf_full_lambda(qidx, 32, 8)), 8, 32, lm)
    cfg[0]['keep_levels'] = 1
    md = wf.WavefrontMD(gpu, cfg)
    md.upload(img[None])
    streams = check_streams(gpu, oracle, md, [img], 8)
    md.close()
    # every leaf size of the tree appears in the stream
    _, _, blocks, _ = oracle.av1_decode_still(streams[0], 512, 512, 8)
    assert len(set(int(s) for s in blocks['size'])) >= 3


def test_sb128_tiles(gpu, oracle):
    img = synth.config4_tile_column(256, 512, seed=4100)
    qidx = synth.qp_to_qindex(28)
    cfg = wf.funnel_cfg(512, 256, 10, qidx, int(gpu.lib().zav1_kf_full_lambda(qidx, 28, 10)), 6, 28, sb=128,
                        tiles=(2, 2), leaf=32)
    cfg[0]['keep_levels'] = 1
    md = wf.WavefrontMD(gpu, cfg)
    md.upload(img[None])
    check_streams(gpu, oracle, md, [img], 10)
    md.close()
