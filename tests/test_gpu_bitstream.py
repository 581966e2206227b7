"""End to end on the GPU: the wavefront mode decision (keep_levels) ->
zav1_wf_write_stills (host AV1 writer) -> the oracle's decoder written from
the specification.  The decoded frame must equal the device reconstruction
sample for sample and the decoded syntax the device decisions (mode, filter
intra, tx type, eob per leaf): the stream is a valid AV1 monochrome still of
exactly what the mode decision chose.  And the streams are byte-identical
to the ones the same writer makes from the oracle's sequential mode
decision (or_wf_frame_levels) of the same images.  Uniform 16x16 leaves at
10 bits (preset 4 funnel, 2 images), the PD0 trees at preset 8 (config 1),
SB128 with 2x2 tiles at preset 6 and 32x32 leaves."""
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


import test_bitstream as tb  # noqa: E402

wf = _mod('wavefront')
part = _mod('partition')
synth = _mod('synth')


def oracle_streams(z, O, cfg, images, leaf_map):
    c = cfg[0]
    W, H, bd, sb, leaf = int(c['width']), int(c['height']), int(c['bd']), int(c['sb']), int(c['leaf'])
    tiles = (int(c['tile_cols']), int(c['tile_rows']))
    _, dec, lev = O.wf_frame(cfg, np.stack(images), leaf_map, levels=True)
    out = []
    for i in range(len(images)):
        lm = None if leaf_map is None else leaf_map[i * (W // 64) * (H // 64) * 256:(i + 1) * (W // 64) * (H // 64) * 256]
        leaves, levels = tb.leaves_of_decisions(dec[i], lev[i], leaf_map is not None, leaf, lm, c['cands'])
        rc, s = tb.write(z, W, H, bd, sb, tiles, int(c['qindex']), leaves, levels)
        assert rc == 0, rc
        out.append(s)
    return out


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
    streams = check_streams(gpu, oracle, md, imgs, 10)
    # the plain per-call form keeps the same levels: cleared, every leaf
    # redone one call at a time, the streams come back byte-identical
    gpu.check(gpu.lib().zav1_memset(md.info.d_levels, 0, 4 * md.info.n_levels, None), 'memset')
    md.set_raw_decisions(np.zeros_like(md.raw_decisions()))
    md.run_plain(0, md.info.n_leaves)
    gpu.sync()
    assert md.write_stills(0, 2, threads=2) == streams
    md.close()
    assert streams == oracle_streams(gpu, oracle, cfg, imgs, None)


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
    assert streams == oracle_streams(gpu, oracle, cfg, [img], lm)
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
    streams = check_streams(gpu, oracle, md, [img], 10)
    md.close()
    assert streams == oracle_streams(gpu, oracle, cfg, [img], None)


def test_pd0_trees_10bit_two_images_graph(gpu, oracle):
    """The preset-8 path on two 10-bit images at once: the device PD0 maps
    (every leaf size from 8x8 to 64x64, so a layer's sizes run as parallel
    stream branches), the plan's decision eager and as a graph, the streams
    -- decoded to the device's reconstruction and byte-identical to the
    streams of the oracle's decisions over the same maps."""
    imgs = [synth.config2_luma(256, 256, seed=2400 + k) for k in range(2)]
    qidx = synth.qp_to_qindex(40)
    pc = part.pd0_config(gpu, 256, 256, 10, 40, qidx, 8, n_images=2, tiles=(2, 2))  # 128-pixel tiles
    d = gpu.DeviceArray.from_host(np.ascontiguousarray(np.concatenate(imgs, axis=0)))
    pp = part.PartitionPass(gpu, pc)
    pp.run(d)
    gpu.sync()
    lm = np.ascontiguousarray(pp.results()[2].reshape(-1), np.uint8)
    pp.close()
    d.free()
    sizes = set(int(v) for v in np.unique(lm))
    print('leaf sizes (log2):', sorted(sizes), flush=True)
    assert len(sizes) >= 3
    cfg = wf.tree_cfg(256, 256, 10, qidx, int(gpu.lib().zav1_kf_full_lambda(qidx, 40, 10)), 8, 40, lm, n_images=2,
                      tiles=(2, 2))
    cfg[0]['keep_levels'] = 1
    md = wf.WavefrontMD(gpu, cfg)
    md.upload(np.stack(imgs))
    streams = check_streams(gpu, oracle, md, imgs, 10)
    st = gpu.stream_create()
    md.run(True, st)
    gpu.stream_sync(st)
    assert md.write_stills(0, 2, threads=2, stream=st) == streams
    gpu.stream_destroy(st)
    md.close()
    assert streams == oracle_streams(gpu, oracle, cfg, imgs, lm)
