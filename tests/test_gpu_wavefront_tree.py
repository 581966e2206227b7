"""The wavefront mode decision over the partition pre-pass's trees on the
GPU: the leaves of the PD0 leaf maps (mixed 8x8 .. 64x64 leaves), scheduled
by their coding-order dependencies, against the oracle's sequential funnel
over the same map (or_wf_frame with a partition map), bit-exact in every
decision record and the reconstruction; graph == eager; the plain per-call
form on the first leaves == the batched form.

Config 1 end to end: 512x512 8-bit, qp 32, preset 8 -- the PD0 pass (LVL_1,
variance cap, small-eob rate) decides the tree, the preset-8 funnel decides
its leaves: the reference's luma mode decision at preset 8 (pred-depth-only,
pipeline.rs:8000-8060).  Plus a 10-bit frame at preset 6 with 2x2 tiles and
a hand-made map with every leaf size, SB128."""
import importlib.util
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIELDS = ('cand', 'tx_type', 'mode', 'fi_mode', 'eob', 'rate', 'mode_rate', 'dist', 'cost')


def _mod(name):
    path = os.path.join(ROOT, 'zenav1-svt_amd', name + '.py')
    spec = importlib.util.spec_from_file_location('zav1_' + name, path)
    m = importlib.util.module_from_spec(spec)
    sys.modules['zav1_' + name] = m
    spec.loader.exec_module(m)
    return m


part = _mod('partition')
wf = _mod('wavefront')
synth = _mod('synth')


def pd0_maps(z, images, bd, qp, preset, tiles=(0, 0)):
    h, w = images[0].shape
    cfg = part.pd0_config(z, w, h, bd, qp, synth.qp_to_qindex(qp), preset, n_images=len(images), tiles=tiles)
    d = z.DeviceArray.from_host(np.ascontiguousarray(np.concatenate(images, axis=0)))
    pp = part.PartitionPass(z, cfg)
    pp.run(d)
    z.sync()
    _, _, lm = pp.results()
    pp.close()
    d.free()
    return np.ascontiguousarray(lm.reshape(-1), np.uint8)


def decide(z, O, images, bd, qp, preset, leaf_map, sb=64, tiles=(1, 1), graph=True, plain=0):
    h, w = images[0].shape
    qidx = synth.qp_to_qindex(qp)
    lam = int(z.lib().zav1_kf_full_lambda(qidx, qp, bd))
    cfg = wf.tree_cfg(w, h, bd, qidx, lam, preset, qp, leaf_map, n_images=len(images), sb=sb, tiles=tiles)
    print('tree wavefront: %dx%d x%d bd %d qp %d preset %d sb %d tiles %s' % (w, h, len(images), bd, qp, preset,
                                                                           sb, tiles), flush=True)
    md = wf.WavefrontMD(z, cfg)
    md.upload(np.stack(images))
    st = z.stream_create()
    md.run(False, st)
    z.stream_sync(st)
    rec, dec = md.recon(), md.decisions_cells()
    out = {'steps': md.info.n_steps, 'leaves': md.info.n_leaves}
    if graph:
        md.run(True, st)
        z.stream_sync(st)
        out['graph_same'] = bool(np.array_equal(md.recon(), rec) and np.array_equal(md.decisions_cells(), dec))
    if plain:
        raw = md.raw_decisions()
        md.set_raw_decisions(np.zeros_like(raw))
        md.run_plain(0, plain, st)
        z.stream_sync(st)
        out['plain_same'] = bool(np.array_equal(md.raw_decisions()[:plain], raw[:plain]))
    md.close()
    z.stream_destroy(st)
    ref_rec, ref_dec = O.wf_frame(cfg, np.stack(images), leaf_map)
    assert np.array_equal(rec, ref_rec), 'reconstruction differs from the oracle'
    for f in FIELDS:
        assert np.array_equal(dec[f], ref_dec[f]), 'decision field %s differs from the oracle' % f
    return out


def _unit_sizes(lm):
    return sorted(set(int(v) for v in np.unique(lm)))


def test_config1_pd0_tree_preset8(gpu, oracle):
    img = synth.config1_luma(512, 512)
    lm = pd0_maps(gpu, [img], 8, 32, 8)
    # the device map equals the oracle pass's
    _, olm, _ = oracle.pd0_frame(img, oracle.pd0_cfg(1, 32, synth.qp_to_qindex(32), cap_max_block=1,
                                                   coeff_rate_est_lvl=2))
    assert np.array_equal(lm, olm.reshape(-1))
    assert len(_unit_sizes(lm)) >= 2
    out = decide(gpu, oracle, [img], 8, 32, 8, lm, plain=24)
    assert out['graph_same'] and out['plain_same']


def test_10bit_preset6_tiles(gpu, oracle):
    imgs = [synth.config2_luma(256, 256, seed=2100 + k) for k in range(2)]
    lm = pd0_maps(gpu, imgs, 10, 40, 6, tiles=(2, 2))
    out = decide(gpu, oracle, imgs, 10, 40, 6, lm, tiles=(2, 2))
    assert out['graph_same']


def test_every_leaf_size_sb128(gpu, oracle):
    """A hand-made map with 64, 32, 16 and 8 leaves in every unit, SB128 (four
    units per superblock in Z-order), 8-bit and 10-bit."""
    rng = np.random.default_rng(11)
    units = []
    for u in range(16):
        m = np.zeros((16, 16), np.uint8)

        def go(sq, x, y):
            if sq == 8 or (sq < 64 or u % 5 == 0) and rng.random() < 0.35 or (u % 5 == 0 and sq == 64):
                m[y // 4:(y + sq) // 4, x // 4:(x + sq) // 4] = int(np.log2(sq))
                return
            h = sq // 2
            for i in range(4):
                go(h, x + (i & 1) * h, y + (i >> 1) * h)
        go(64, 0, 0)
        units.append(m.ravel())
    lm = np.ascontiguousarray(np.concatenate(units), np.uint8)
    assert _unit_sizes(lm) == [3, 4, 5, 6]
    img8 = synth.config1_luma(256, 256, seed=1003)
    decide(gpu, oracle, [img8], 8, 28, 4, lm, sb=128, graph=False)
    img10 = synth.config2_luma(256, 256, seed=2003)
    decide(gpu, oracle, [img10], 10, 50, 8, lm, sb=128, graph=False)


def test_autotune_keeps_the_decisions(gpu, oracle):
    """The autotuner times every trellis form per leaf size and keeps the
    fastest; the forms give the same levels, so the decisions, the
    reconstruction and a graph captured after tuning equal the untuned
    run's (config-2 content, 10 bits, PD0 tree at preset 4: four sizes)."""
    imgs = [synth.config2_luma(256, 256, seed=2400 + k) for k in range(2)]
    lm = pd0_maps(gpu, imgs, 10, 28, 4)
    qidx = synth.qp_to_qindex(28)
    cfg = wf.tree_cfg(256, 256, 10, qidx, int(gpu.lib().zav1_kf_full_lambda(qidx, 28, 10)), 4, 28, lm, n_images=2)
    md = wf.WavefrontMD(gpu, cfg)
    md.upload(np.stack(imgs))
    st = gpu.stream_create()
    md.run(True, st)
    gpu.stream_sync(st)
    rec, dec = md.recon(), md.raw_decisions()
    tuned = md.autotune(0, md.info.n_steps, st)
    print('autotune:', tuned, flush=True)
    assert tuned['forms'] and all(f in wf.WavefrontMD.FORMS for f in tuned['forms'].values())
    assert np.array_equal(md.recon(), rec) and np.array_equal(md.raw_decisions(), dec)
    md.run(True, st)   # a new graph with the tuned forms
    gpu.stream_sync(st)
    assert np.array_equal(md.recon(), rec) and np.array_equal(md.raw_decisions(), dec)
    md.run(False, st)
    gpu.stream_sync(st)
    assert np.array_equal(md.recon(), rec) and np.array_equal(md.raw_decisions(), dec)
    md.close()
    gpu.stream_destroy(st)
