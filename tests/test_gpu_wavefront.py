"""GPU parity of the superblock-wavefront mode decision (zav1_wf_*) with the
oracle's sequential restatement of the leaf funnel in the reference's coding
order: every leaf's decision (candidate, tx type, eob, rates, distortion,
cost) and the whole reconstruction bit-exact, at 8 and 10 bits, leaf sizes
8 / 16 / 32, several images per batch, tiles, partial superblocks, every
rdoq level; the graph replay and the plain per-call form equal the eager
launches."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _wf():
    spec = importlib.util.spec_from_file_location('zav1_wavefront', os.path.join(ROOT, 'zenav1-svt_amd', 'wavefront.py'))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def content(rng, n, H, W, bd):
    """Smooth gradients, a few hard-edged rectangles, noise (the config-1
    generator's ingredients at small size)."""
    mx = (1 << bd) - 1
    yy, xx = np.mgrid[0:H, 0:W]
    out = []
    for _ in range(n):
        a, b = rng.uniform(-1, 1, 2)
        img = (0.5 + 0.35 * np.sin(xx * a / 9.0 + yy * b / 7.0)) * mx
        for _ in range(4):
            x0, y0 = rng.integers(0, W - 8), rng.integers(0, H - 8)
            img[y0:y0 + rng.integers(4, 24), x0:x0 + rng.integers(4, 24)] = rng.uniform(0.1, 0.9) * mx
        img += rng.normal(0, 2 << (bd - 8), (H, W))
        out.append(np.clip(img, 0, mx))
    return np.array(out).astype(np.uint8 if bd == 8 else np.uint16)


CASES = [
    # (W, H, bd, leaf, n_images, tiles, k1, k3, rdoq, qindex, satd_th)
    (128, 128, 8, 16, 1, (1, 1), 8, 2, 2, 120, 10),
    (96, 160, 10, 8, 2, (2, 1), 6, 3, 0, 80, 5),
    (192, 64, 10, 32, 1, (1, 1), 10, 2, 3, 200, 0),
    (80, 48, 8, 16, 3, (1, 1), 4, 4, 1, 40, 20),
    (128, 128, 10, 16, 1, (2, 2), 15, 4, 1, 140, 12),
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_wavefront_matches_oracle(gpu, oracle, case):
    z = gpu
    wf = _wf()
    W, H, bd, leaf, n, tiles, k1, k3, rdoq, qidx, sth = CASES[case]
    rng = np.random.default_rng(case)
    imgs = content(rng, n, H, W, bd)
    lam = int(rng.integers(20000, 2000000))
    cfg = wf.make_cfg(W, H, bd, qidx, lam, n_images=n, leaf=leaf, tiles=tiles, k1=k1, k3=k3, rdoq_level=rdoq,
                      txt_satd_th=sth, cands=wf.preset_candidates(fi=leaf <= 32))
    md = wf.WavefrontMD(z, cfg)
    try:
        md.upload(imgs)
        md.run(use_graph=False)
        z.sync()
        rec, dec = md.recon(), md.decisions()
        orec, odec = oracle.wf_frame(cfg, imgs)
        for f in ('cand', 'tx_type', 'eob', 'rate', 'mode_rate', 'dist', 'cost'):
            bad = np.argwhere(dec[f] != odec[f])
            assert bad.size == 0, (case, f, bad[:3].tolist(), dec[f][tuple(bad[0])], odec[f][tuple(bad[0])])
        assert np.array_equal(rec, orec), case
        # not every leaf took DCT_DCT or DC: the search is live
        assert len(np.unique(dec['cand'])) > 3
    finally:
        md.close()


def test_wavefront_graph_and_plain_equal_eager(gpu):
    z = gpu
    wf = _wf()
    rng = np.random.default_rng(77)
    W, H, bd = 128, 96, 10
    imgs = content(rng, 2, H, W, bd)
    cfg = wf.make_cfg(W, H, bd, 110, 300000, n_images=2, leaf=16, tiles=(1, 1), k1=6, k3=2, rdoq_level=2,
                      txt_satd_th=10)
    outs, runs = [], {}
    for mode in ('eager', 'graph', 'graph2', 'plain'):
        md = wf.WavefrontMD(z, cfg)
        try:
            md.upload(imgs)
            if mode == 'eager':
                md.run(False)
            elif mode.startswith('graph'):
                st = z.stream_create()
                md.run(True, st)
                if mode == 'graph2':       # a second replay over the same inputs
                    md.run(True, st)
                z.stream_sync(st)
                z.stream_destroy(st)
            else:
                md.run_plain(0, md.info.n_leaves)
            z.sync()
            outs.append((md.recon(), md.decisions()))
            c = np.zeros(1, np.uint64)
            z.check(z.lib().zav1_memcpy_d2h(c.ctypes.data, md.info.d_trellis_runs, 8, None), 'runs')
            runs[mode] = int(c[0])
        finally:
            md.close()
    for r, d in outs[1:]:
        assert np.array_equal(r, outs[0][0]) and np.array_equal(d, outs[0][1])
    # the trellis-run counter: the same per run, summed over replays, within
    # the MDS3 jobs (leaves x survivors x types); the plain form does not count
    jobs = md.info.n_leaves * 2 * md.info.n_types
    assert 0 < runs['eager'] <= jobs and runs['graph'] == runs['eager'] and runs['graph2'] == 2 * runs['eager']
    assert runs['plain'] == 0


def test_wavefront_rejects_bad_configs(gpu):
    z = gpu
    wf = _wf()
    for kw in (dict(leaf=12), dict(k1=2, k3=3), dict(rdoq_level=4), dict(sb=32)):
        args = dict(n_images=1, leaf=16, k1=6, k3=2, rdoq_level=1)
        args.update(kw)
        cfg = wf.make_cfg(64, 64, 8, 100, 1000, **args)
        with pytest.raises(z.Zav1Error):
            wf.WavefrontMD(z, cfg)
