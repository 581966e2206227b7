"""The partition pre-pass on the GPU (zav1_pd0_run) against the oracle's
sequential restatement (or_pd0_frame), bit-exact: per-unit variance maps,
levels, depth caps and subres decisions, every node record (visited /
tested / split flags and the tested blocks' costs) and the leaf maps.

Cases: the reference's gradient-64 unit at its captured qps for every level
(whose trees the reference pins, pd0.rs:2700-2780); config 1's 512x512
8-bit content at preset 8 (LVL_1 with the variance cap and the small-eob
rate); config 2's 10-bit content MSB-truncated at preset 4 (LVL_1, exact
rate) and at preset 9 (LVL_0); eff-M9 8-bit content (LVL_5 with subres
units and LVL_6 units in one frame); preset 2's 4x4 blocks; two stacked
images with tiles; uniform and extreme content."""
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


part = _mod('partition')
synth = _mod('synth')


def gradient64():
    r = np.arange(64)[:, None]
    c = np.arange(64)[None, :]
    return (((r * 255) // 64) ^ ((c * 3) & 0x3f)).astype(np.uint8)


def _oracle_cfg(O, c):
    oc = O.pd0_cfg(int(c['level']), int(c['qp']), int(c['qindex']), ires=int(c['ires']),
                   cap_max_block=int(c['cap_max_block']), max_tx=int(c['max_tx']), min_sq=int(c['min_sq']),
                   coeff_rate_est_lvl=int(c['coeff_rate_est_lvl']), tile_w=int(c['tile_w']),
                   tile_h=int(c['tile_h']), lam=int(c['lambda']))
    return oc


def run_and_check(z, O, images, bd, qp, preset, tiles=(0, 0), max_tx=64):
    """images: list of equal-size planes (u8 at bd 8, u16 at bd 10)."""
    h, w = images[0].shape
    qidx = synth.qp_to_qindex(qp)
    cfg = part.pd0_config(z, w, h, bd, qp, qidx, preset, n_images=len(images), tiles=tiles, max_tx=max_tx)
    stacked = np.ascontiguousarray(np.concatenate(images, axis=0))
    print('pd0 case: %dx%d x%d bd %d qp %d preset %d level %d' % (w, h, len(images), bd, qp, preset,
                                                                 int(cfg[0]['level'])), flush=True)
    d_plane = z.DeviceArray.from_host(stacked)
    pp = part.PartitionPass(z, cfg)
    pp.run(d_plane)
    z.sync()
    sb, nodes, lm = pp.results()
    pp.close()
    d_plane.free()
    oc = _oracle_cfg(O, cfg[0])
    per = (w // 64) * (h // 64)
    for i, img in enumerate(images):
        p8 = (img >> (bd - 8)).astype(np.uint8)
        on, olm, omodes = O.pd0_frame(p8, oc)
        s = slice(i * per, (i + 1) * per)
        assert np.array_equal(sb['mode'][s], omodes), 'levels differ (image %d)' % i
        for u in range(per):
            exp_var = O.b64_variance(p8, (u % (w // 64)) * 64, (u // (w // 64)) * 64)
            assert np.array_equal(sb['var'][i * per + u], exp_var), (i, u)
        nps = nodes.shape[1]
        assert np.array_equal(nodes['flags'][s], on['flags'][:, :nps]), 'node flags differ (image %d)' % i
        bad = np.nonzero(nodes['cost'][s] != on['cost'][:, :nps])
        assert bad[0].size == 0, 'node costs differ (image %d): first at unit %d node %d: %d vs %d' % (
            i, bad[0][0], bad[1][0], nodes['cost'][s][bad[0][0], bad[1][0]], on['cost'][bad[0][0], bad[1][0]])
        assert np.array_equal(lm[s], olm), 'leaf maps differ (image %d)' % i
    return sb, nodes, lm


@pytest.mark.parametrize('preset,qp', [(9, 20), (9, 40), (9, 55), (8, 20), (6, 40), (4, 55)])
def test_gradient64_every_level(gpu, oracle, preset, qp):
    sb, nodes, lm = run_and_check(gpu, oracle, [gradient64()], 8, qp, preset)
    # the reference's captured trees (pd0.rs:2700, :2770)
    expect = {(9, 20): [16] * 16, (9, 40): [32] * 4, (9, 55): [64], (6, 40): [32] * 4, (4, 55): [64]}
    if (preset, qp) in expect:
        assert part.leaf_sizes(lm[0]) == expect[(preset, qp)]


def test_gradient64_10bit_level0(gpu, oracle):
    y10 = gradient64().astype(np.uint16) << 2 | 3   # MSB-truncation drops the low bits
    sb, nodes, lm = run_and_check(gpu, oracle, [y10], 10, 20, 10)
    assert part.leaf_sizes(lm[0]) == [32] * 4           # pd0.rs:2347 (LVL_0, q20)
    assert nodes['cost'][0][1:5].tolist() == [26185862, 28871046, 22111622, 22521222]


def test_config1_preset8(gpu, oracle):
    run_and_check(gpu, oracle, [synth.config1_luma(512, 512)], 8, 32, 8)


@pytest.mark.parametrize('preset', [4, 9])
def test_config2_crop_10bit(gpu, oracle, preset):
    img = synth.config2_luma(512, 512, seed=2000)
    run_and_check(gpu, oracle, [img], 10, 40, preset)


def test_eff_m9_mixed_levels_and_subres(gpu, oracle):
    rng = np.random.default_rng(7)
    # smooth ramps (detector demotes, subres passes) beside noisy units (LVL_6)
    img = np.zeros((256, 256), np.uint8)
    ramp = (np.arange(256)[None, :] // 2 + np.arange(256)[:, None] // 4).astype(np.uint8)
    img[:] = ramp
    img[64:128, :] = rng.integers(0, 256, (64, 256))
    img[192:, 128:] = rng.integers(100, 140, (64, 128))
    for qp in (30, 55):
        sb, _, _ = run_and_check(gpu, oracle, [img], 8, qp, 10)
        # both levels, and the subres transforms of the units whose 64x64
        # passes the odd/even check (12 of 16 here)
        assert set(np.unique(sb['mode'])) >= {5, 6}
        assert int(sb['subres'].sum()) == 12


def test_preset2_4x4_blocks_two_images_tiles(gpu, oracle):
    a = synth.config1_luma(256, 256, seed=1001)
    b = synth.config1_luma(256, 256, seed=1002)
    sb, nodes, lm = run_and_check(gpu, oracle, [a, b], 8, 28, 2, tiles=(2, 2))
    assert nodes.shape[1] == 341


def test_uniform_and_extreme_content(gpu, oracle):
    run_and_check(gpu, oracle, [np.full((128, 128), 128, np.uint8), np.zeros((128, 128), np.uint8)], 8, 40, 6)
    chk = ((np.indices((128, 128)).sum(axis=0) & 1) * 255).astype(np.uint8)
    run_and_check(gpu, oracle, [chk], 8, 20, 8)
    run_and_check(gpu, oracle, [chk], 8, 63, 10)
    run_and_check(gpu, oracle, [chk.astype(np.uint16) * 4 + 3], 10, 50, 4, max_tx=32)
