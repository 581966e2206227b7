"""The AV1 still writer (zav1_write_still, host) against the oracle's
decoder (or_av1_decode_still, written from the specification's decoding
process): random leaves -- every leaf size 8..64 in random quadtrees, every
intra mode and angle delta, filter intra, every tx type of each size's intra
set, random levels with tight eobs, large levels through the Golomb
remainder, skipped leaves -- written and decoded back; the decoded syntax
must equal the written one exactly (positions, sizes, modes, deltas,
filter-intra modes, skip, tx types, eobs and every level), at 8 and 10 bits,
SB64 and SB128, one and several tiles.  Also the writer's refusals.

The oracle's own mode decision (or_wf_frame_levels: uniform leaves and a
partition map) written by the library's writer decodes to exactly the
oracle's reconstruction: the stream carries what the funnel decided, and
the decoder's prediction and inverse transforms (the specification's)
agree with the funnel's.  The same check over the device's decisions runs
on the GPU in test_gpu_bitstream.py."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod(name):
    path = os.path.join(ROOT, 'zenav1-svt_amd', name + '.py')
    spec = importlib.util.spec_from_file_location('zav1_' + name, path)
    m = importlib.util.module_from_spec(spec)
    sys.modules['zav1_' + name] = m
    spec.loader.exec_module(m)
    return m


bs = _mod('bitstream')
STILL_LEAF = bs.STILL_LEAF
SET7 = [0, 1, 2, 3, 9, 10, 11]
SET5 = [0, 1, 2, 3, 9]


def scan_of(oracle, size, tx_type):
    txs = {4: 0, 8: 1, 16: 2, 32: 3, 64: 4}[size]
    cls = 0 if tx_type < 10 else (2 if tx_type & 1 else 1)
    return oracle.scan(txs, cls)


def random_still(oracle, rng, W, H, sb, big=False):
    leaves, levels = [], []
    off = 0

    def leaf(x, y, s):
        nonlocal off
        n = min(s, 32)
        mode = int(rng.integers(0, 13))
        delta = int(rng.integers(-3, 4)) if 1 <= mode <= 8 else 0
        fi = int(rng.integers(0, 5)) if (mode == 0 and s <= 32 and rng.random() < 0.4) else 5
        skip = rng.random() < 0.2
        if s >= 32:
            tt = 0
        else:
            tt = int(rng.choice(SET5 if s == 16 else SET7))
        q = np.zeros(n * n, np.int32)
        eob = 0
        if not skip:
            eob = int(rng.integers(1, n * n + 1)) if rng.random() < 0.3 else int(rng.integers(1, min(n * n, 40) + 1))
            sc = scan_of(oracle, s, tt)
            for c in range(eob):
                if rng.random() < 0.5 or c == eob - 1:
                    mag = int(rng.integers(1, 4)) if rng.random() < 0.8 else int(rng.integers(3, 20))
                    if big and rng.random() < 0.05:
                        mag = int(rng.integers(15, 70000))
                    q[sc[c]] = mag if rng.random() < 0.5 else -mag
        leaves.append((x, y, s, mode, delta, fi, tt, eob, off))
        levels.append(q)
        off += q.size

    def walk(x, y, s):
        if x >= W or y >= H:
            return
        if s <= 64 and (s == 8 or rng.random() < 0.35):
            leaf(x, y, s)
            return
        h = s // 2
        for i in range(4):
            walk(x + (i & 1) * h, y + (i >> 1) * h, h)

    for y in range(0, H, sb):
        for x in range(0, W, sb):
            walk(x, y, sb)
    arr = np.zeros(len(leaves), STILL_LEAF)
    for i, (x, y, s, mode, delta, fi, tt, eob, o) in enumerate(leaves):
        arr[i] = (x, y, s, mode, delta, fi, tt, (0, 0, 0), eob, 0, o)
    return arr, np.concatenate(levels)


def write(zav1, W, H, bd, sb, tiles, qidx, leaves, levels, threads=0):
    return bs.write_still(zav1, W, H, bd, sb, tiles, qidx, leaves, levels, threads=threads)


@pytest.mark.parametrize('W,H,bd,sb,tiles,qidx,seed', [
    (128, 128, 8, 64, (1, 1), 100, 1),
    (256, 192, 10, 64, (2, 1), 180, 2),
    (256, 256, 8, 128, (1, 1), 20, 3),
    (512, 256, 10, 64, (4, 2), 60, 4),
    (256, 256, 8, 64, (2, 2), 255, 5),
])
def test_written_syntax_decodes_back(zav1, oracle, W, H, bd, sb, tiles, qidx, seed):
    rng = np.random.default_rng(seed)
    leaves, levels = random_still(oracle, rng, W, H, sb, big=True)
    rc, stream = write(zav1, W, H, bd, sb, tiles, qidx, leaves, levels)
    assert rc == 0, rc
    frame, info, blocks, dlev = oracle.av1_decode_still(stream, W, H, bd)
    assert (info.width, info.height, info.bd, info.sb, info.qindex) == (W, H, bd, sb, qidx)
    assert (info.tile_cols, info.tile_rows) == tiles
    assert info.n_blocks == len(leaves)
    by_pos = {(int(l['x']), int(l['y'])): i for i, l in enumerate(leaves)}
    for b in blocks:
        i = by_pos[(int(b['x']), int(b['y']))]
        l = leaves[i]
        assert (b['size'], b['mode'], b['delta'], b['fi_mode']) == (l['size'], l['mode'], l['delta'], l['fi_mode'])
        assert b['eob'] == l['eob'] and bool(b['skip']) == (l['eob'] == 0)
        if l['eob']:
            assert b['tx_type'] == l['tx_type']
            n = min(int(l['size']), 32) ** 2
            assert np.array_equal(dlev[b['level_off']:b['level_off'] + n], levels[l['level_off']:l['level_off'] + n])


def test_threads_do_not_change_the_stream(zav1, oracle):
    rng = np.random.default_rng(9)
    leaves, levels = random_still(oracle, rng, 512, 256, 64)
    s1 = write(zav1, 512, 256, 8, 64, (4, 2), 90, leaves, levels, threads=1)[1]
    s8 = write(zav1, 512, 256, 8, 64, (4, 2), 90, leaves, levels, threads=8)[1]
    assert s1 == s8


def test_writer_refuses_bad_input(zav1, oracle):
    rng = np.random.default_rng(3)
    leaves, levels = random_still(oracle, rng, 128, 128, 64)
    assert write(zav1, 128, 128, 8, 64, (3, 1), 100, leaves, levels)[0] == zav1.ZAV1_EINVAL   # not a power of two
    assert write(zav1, 128, 128, 8, 64, (1, 1), 0, leaves, levels)[0] == zav1.ZAV1_EINVAL     # lossless
    gap = leaves[1:].copy()
    assert write(zav1, 128, 128, 8, 64, (1, 1), 100, gap, levels)[0] == zav1.ZAV1_EINVAL     # uncovered cells
    assert 'cells no leaf covers' in zav1.last_error()
    loose = leaves.copy()
    k = int(np.nonzero(loose['eob'] > 1)[0][0])
    sc = scan_of(oracle, int(loose[k]['size']), int(loose[k]['tx_type']))
    lv2 = levels.copy()
    lv2[int(loose[k]['level_off']) + int(sc[int(loose[k]['eob']) - 1])] = 0
    assert write(zav1, 128, 128, 8, 64, (1, 1), 100, loose, lv2)[0] == zav1.ZAV1_EINVAL       # eob not tight
    assert 'last nonzero is not at the eob' in zav1.last_error()


def leaves_of_decisions(dec, lev, cells, leaf=None, leaf_map=None, cands=None):
    """Still leaves (raster order of their top-left cells, not coding order)
    from one image's oracle decisions and level plane; the angle delta is the
    winning candidate's (cands: the configuration's candidate records)."""
    H, W = lev.shape
    out, chunks, off = [], [], 0
    if leaf_map is not None:
        uw = W // 64
        size_at = lambda x, y: 1 << int(leaf_map[((y // 64) * uw + x // 64) * 256 + ((y % 64) // 4) * 16 + (x % 64) // 4])
    for (r, c), d in np.ndenumerate(dec):
        g = 8 if cells else leaf
        x, y = c * g, r * g
        if cells:
            s = size_at(x, y)
            if x % s or y % s:
                continue
        else:
            s = leaf
        p = min(s, 32)
        q = np.ascontiguousarray(lev[y:y + p, x:x + p]).ravel()
        delta = int(cands[d['cand']]['delta'])
        out.append((x, y, s, d['mode'], delta, d['fi_mode'], d['tx_type'], (0, 0, 0), d['eob'], 0, off))
        chunks.append(q)
        off += q.size
    arr = np.zeros(len(out), STILL_LEAF)
    for i, t in enumerate(out):
        arr[i] = t
    return arr, np.concatenate(chunks)


@pytest.mark.parametrize('case', ['uniform16_10bit_tiles', 'map_8bit', 'uniform32_sb128', 'preset4_angles'])
def test_oracle_decisions_decode_to_the_reconstruction(zav1, oracle, case):
    wf = _mod('wavefront')
    synth = _mod('synth')
    if case == 'uniform16_10bit_tiles':
        W, H, bd, qp, preset, sb, tiles, leaf = 128, 128, 10, 40, 8, 64, (2, 2), 16
        img = synth.config2_luma(H, W, seed=2300)
    elif case == 'uniform32_sb128':
        W, H, bd, qp, preset, sb, tiles, leaf = 256, 128, 8, 28, 8, 128, (1, 1), 32
        img = synth.config1_luma(H, W, seed=1300)
    elif case == 'preset4_angles':
        W, H, bd, qp, preset, sb, tiles, leaf = 128, 128, 10, 28, 4, 64, (1, 1), 8
        img = synth.config2_luma(H, W, seed=2301)
    else:
        W, H, bd, qp, preset, sb, tiles, leaf = 128, 128, 8, 32, 8, 64, (1, 1), 16
        img = synth.config1_luma(H, W, seed=1301)
    qidx = synth.qp_to_qindex(qp)
    lam = int(zav1.lib().zav1_kf_full_lambda(qidx, qp, bd))
    lm = None
    if case == 'map_8bit':
        _, lm, _ = oracle.pd0_frame(img, oracle.pd0_cfg(1, qp, qidx, cap_max_block=1, coeff_rate_est_lvl=2))
        lm = np.ascontiguousarray(lm.reshape(-1), np.uint8)
        assert len(set(lm.tolist())) >= 2
        cfg = wf.tree_cfg(W, H, bd, qidx, lam, preset, qp, lm, sb=sb, tiles=tiles)
    else:
        cfg = wf.funnel_cfg(W, H, bd, qidx, lam, preset, qp, sb=sb, tiles=tiles, leaf=leaf)
    rec, dec, lev = oracle.wf_frame(cfg, img[None], lm, levels=True)
    leaves, levels = leaves_of_decisions(dec[0], lev[0], lm is not None, leaf, lm, cfg[0]['cands'])
    # the tight eob of each leaf's levels is its decided eob
    rc, stream = write(zav1, W, H, bd, sb, tiles, qidx, leaves, levels)
    assert rc == 0, rc
    frame, info, blocks, _ = oracle.av1_decode_still(stream, W, H, bd)
    assert info.n_blocks == len(leaves)
    assert np.array_equal(frame, rec[0]), 'decoded frame differs from the funnel reconstruction'
    if lm is None:  # the package's vectorised helper builds the same stream
        ul, uv = bs.leaves_uniform(dec[0], lev[0], leaf, cfg[0]['cands'])
        assert write(zav1, W, H, bd, sb, tiles, qidx, ul, uv) == (0, stream)
    if case == 'preset4_angles':  # nonzero angle deltas and filter intra were decided and coded
        assert (leaves['delta'] != 0).any() and (leaves['fi_mode'] < 5).any()
