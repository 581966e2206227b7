"""The oracle against the C encoder's own full-loop records.

The reference holds instrumented captures of C SVT-AV1 encoding 8-bit
gradient stills at preset 6 (rust/docs/captures/gradient_*_p6.m6fnl.txt,
fixture tests/golden/funnel_full.json by make_funnel_golden.py).  At the
frame-origin block no neighbour is coded, so every prediction follows from
the source alone and each captured number is a function of this repository's
oracle chain:

  MDS1 (leaf_funnel/mod.rs:3067-3110): DCT_DCT at tx depth 0, quantize_b, no
    RDOQ, frequency-domain distortion (sse_i32 + the 64-fold's discarded
    energy, shifted by (1 - log_scale) * 2), coefficient rate cost_coeffs_txb
    (leaf_funnel/coeff_rate.rs:60) with the default-CDF tables of the frame's
    qindex (enc/quant.rs:514)  ->  `ybits`, `ydist`, `eob0`;
  MDS3 tx search (leaf_funnel/mod.rs:7140-7330): quantize_fp + optimize_b
    (enc/quant.rs:1312) at the picture's rdoq level, or quantize_b at level 0,
    spatial distortion (SSE << 4 after the inverse transform and recon), the
    same rate  ->  TXT `eob`, `bits`, `dist` at depth 0 and for the depth-1
    transform block at the origin (txb 0), whose contexts the capture gives.

This pins, end to end on the C encoder's numbers: the cost tables built from
the default CDFs (txb-skip, eob, base, base-eob, br, dc-sign entries), the
coefficient-rate context chain over multi-coefficient blocks, the RDOQ
trellis with real tables and both cut-off levels, the 32x32 / 64x64 / 16x16
forward DCT with the 64-fold, quantize_b / quantize_fp, the inverse
transform, recon and both distortions, and the tx-type rate of 16x16.
"""
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FULL = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'funnel_full.json')))

# Picture rdoq level per capture cell: enc/quant.rs:34-39 lists g64 q20/q40/q55
# and g128 q20 (0, 3, 3, 2); g128 q40 and q55 follow from the same policy
# (derive_intra_coeff_level :1502, rdoq_level_allintra :1535): pic_avg_variance
# 1483 / qp = 37 and 26, below the < 240p VLOW threshold 25 * 1.7 -> level 2.
RDOQ_LEVEL = {('gradient_64x64', 20): 0, ('gradient_64x64', 40): 3, ('gradient_64x64', 55): 3,
              ('gradient_128x128', 20): 2, ('gradient_128x128', 40): 2, ('gradient_128x128', 55): 2}
CUTOFF = {0: None, 1: (0, 0), 2: (80, 100), 3: (60, 100)}     # enc/quant.rs:1551 rdoq_cutoffs
FIMODE_TO_INTRADIR = [0, 1, 2, 6, 0]                          # leaf_funnel/mod.rs:460
TX_OF_SIDE = {16: 2, 32: 3, 64: 4}
TX_SCALE = {2: 0, 3: 1, 4: 2}


def gradient(w, h):
    r, c = np.mgrid[0:h, 0:w]
    return (((r * 255) // h).astype(np.uint8) ^ ((c * 3) & 0x3f).astype(np.uint8)).astype(np.uint8)


def origin_pred(oracle, img, side, mode, fi):
    job = np.zeros(1, oracle.INTRA_JOB)
    job[0] = (0, 0, side, side, mode, 0, fi, 1, 0, 0, 0, 0, 0, 0, 0, 0)
    return oracle.intra_predict(img, 8, job).astype(np.int64)


def eob_of(q, scan):
    nz = np.nonzero(q.ravel()[scan])[0]
    return int(nz[-1]) + 1 if nz.size else 0


def coeff_rate(oracle, tables, q, eob, tx, scan, tsc, dsc, txt_rate):
    txb, eobc = tables
    _, txs_ctx, ems = oracle.rdoq_size_ctx(tx)
    c = txb[txs_ctx * 2]
    if eob == 0:
        return int(c[tsc * 2 + 1])            # txb_skip_cost[ctx][1]
    return oracle.cost_coeffs_txb(q.ravel(), eob, scan, c, eobc[ems, 0], tx, 0, 0, tsc, dsc, txt_rate)


def unit(oracle, img, qidx, side, mode, fi, *, rdoq_level, lam, tsc=0, dsc=0, spatial):
    """One luma transform unit of the capture chain at the origin."""
    tx = TX_OF_SIDE[side]
    pred = origin_pred(oracle, img, side, mode, fi)
    res = img[:side, :side].astype(np.int64) - pred
    coef = oracle.fwd_txfm2d(res.astype(np.int32), tx, 0)
    p = min(side, 32)
    packed = np.ascontiguousarray(coef[:p, :p])
    row = oracle.quant_row(qidx, 8)
    scan = oracle.scan(tx, 0)
    ls = TX_SCALE[tx]
    if rdoq_level == 0:
        q, dq = oracle.quantize('b', packed, row, ls, 8)
        eob = eob_of(q, scan)
    else:
        q, dq = oracle.quantize('fp', packed, row, ls, 8)
        eob = eob_of(q, scan)
        if eob:
            txb, eobc = TABLES[qidx]
            _, txs_ctx, ems = oracle.rdoq_size_ctx(tx)
            rdmult = (lam * 17 + 2) >> 2                  # rdoq_rdmult_full (enc/quant.rs:923), luma
            num, den = CUTOFF[rdoq_level]
            q, dq, eob = oracle.optimize_b(packed.ravel(), q, dq, eob, scan, [row.dequant[0], row.dequant[1]],
                                           txb[txs_ctx * 2], eobc[ems, 0], rdmult, tx, 0, tsc, dsc, 0, num, den)
    txt_rate = oracle.intra_txt_rate(tx, FIMODE_TO_INTRADIR[fi] if fi < 5 else mode, 0)
    bits = coeff_rate(oracle, TABLES[qidx], q, eob, tx, scan, tsc, dsc, txt_rate)
    if spatial:
        full = np.zeros((side, side), np.int32)
        full[:p, :p] = np.asarray(dq).reshape(p, p)
        inv = oracle.inv_txfm2d(full, tx, 0, 8) if eob else np.zeros((side, side), np.int32)
        rec = np.clip(pred + inv, 0, 255)
        dist = int(((img[:side, :side].astype(np.int64) - rec) ** 2).sum()) << 4
    else:
        d = oracle.sse_i32(packed, dq) if eob else oracle.sq_sum_i32(packed)
        d += int((coef.astype(np.int64) ** 2).sum() - (packed.astype(np.int64) ** 2).sum())
        sh = (1 - ls) * 2
        dist = d >> sh if sh >= 0 else d << -sh
    return eob, bits, dist


TABLES = {}


@pytest.fixture(scope='module', autouse=True)
def tables(oracle):
    for qi in (80, 160, 220):
        TABLES[qi] = oracle.build_coeff_costs(qi)


def cells():
    for f in FULL['files']:
        b = f['blocks'][0]
        assert b['org'] == [0, 0]
        yield f, b


@pytest.mark.parametrize('cell', range(6))
def test_mds1_records_at_origin(oracle, cell):
    f, b = list(cells())[cell]
    img = gradient(f['w'], f['h'])
    for r in b['mds1']:
        eob, bits, dist = unit(oracle, img, b['qidx'], b['sq'], r['mode'], r['fi'], rdoq_level=0, lam=b['lambda'],
                               spatial=False)
        assert (eob, bits, dist) == (r['eob0'], r['ybits'], r['ydist']), (f['name'], r)


@pytest.mark.parametrize('cell', range(6))
def test_mds3_txt_records_at_origin(oracle, cell):
    f, b = list(cells())[cell]
    img = gradient(f['w'], f['h'])
    lvl = RDOQ_LEVEL[(f['name'].split('_q')[0], f['qp'])]
    n = 0
    for cand in b['txt']:
        txts = [t for t in cand['records'] if t['kind'] == 'txt']
        for t in txts:
            if t['txb'] != 0 or t['tt'] != 0:
                continue
            side = b['sq'] >> t['dep']
            got = unit(oracle, img, b['qidx'], side, cand['mode'], cand['fi'], rdoq_level=lvl, lam=b['lambda'],
                       tsc=t['tsc'], dsc=t['dsc'], spatial=True)
            assert got == (t['eob'], t['bits'], t['dist']), (f['name'], cand['mode'], cand['fi'], t)
            n += 1
        # depth 0 is the winner's unit when the block keeps depth 0
        d0 = [t for t in txts if t['dep'] == 0][0]
        assert (cand['ybits'], cand['ydist']) == (d0['bits'], d0['dist'])
    assert n == 2 * len(b['txt'])


def test_cost_symbol_closed_form(oracle):
    """av1_cost_symbol = the table lookup of round(-log2(p) * 512) after
    normalising p15 into [2^14, 2^15) and rounding it to 8 bits: within
    512 * log2(1 + 1/256) + 0.5 of the exact cost for every probability,
    exact at powers of two."""
    g = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'coef_cdfs.json')))
    pc = np.array(g['AV1_PROB_COST'])
    i = np.arange(128, 256)
    assert np.array_equal(pc, np.round(-np.log2(i / 256.0) * 512).astype(int))
    for p15 in list(range(1, 40)) + list(range(40, 32768, 97)) + [16384, 32767]:
        exact = -np.log2(p15 / 32768.0) * 512
        assert abs(oracle.cost_symbol(p15) - exact) <= 512 * np.log2(1 + 1 / 256) + 0.5, p15
    for k in range(15):
        assert oracle.cost_symbol(1 << k) == (15 - k) * 512


def test_coeff_cost_tables_structure(oracle):
    """Structural identities of the derived entries (enc/quant.rs:604-633)
    and of probability-derived costs: each binary CDF's two costs describe
    probabilities that sum to one within the cost table's rounding."""
    for qi in (0, 20, 21, 60, 61, 120, 121, 255):
        txb, eob = oracle.build_coeff_costs(qi)
        t = txb.reshape(10, -1)
        base = t[:, 26 + 12:26 + 12 + 336].reshape(10, 42, 8)
        assert np.all(base[:, :, 4] == 0)
        assert np.array_equal(base[:, :, 5], base[:, :, 1] + 512 - base[:, :, 0])
        assert np.array_equal(base[:, :, 6], base[:, :, 2] - base[:, :, 1])
        lps = t[:, 26 + 12 + 336 + 18 + 6:].reshape(10, 21, 26)
        assert np.array_equal(lps[:, :, 13], lps[:, :, 0])
        assert np.array_equal(lps[:, :, 14:], np.diff(lps[:, :, :13], axis=2))
        skip = t[:, :26].reshape(10, 13, 2).astype(float)
        psum = 2 ** (-skip[..., 0] / 512) + 2 ** (-skip[..., 1] / 512)
        assert np.all(np.abs(psum - 1) < 0.02)
    a, b = oracle.build_coeff_costs(61), oracle.build_coeff_costs(120)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))        # one q context
    a = oracle.build_coeff_costs(60)
    assert not np.array_equal(a[0], b[0])


def test_intra_txt_rate_sets(oracle):
    """The intra ext-tx sets (coeff_c.rs:181-212): no type rate at 32 and 64
    points; 16x16 offers DCT, ADST x DCT, ADST_ADST and IDTX; smaller sizes
    add V_DCT / H_DCT; rates are positive exactly on the set's types."""
    allowed16 = {0, 1, 2, 3, 9}
    allowed8 = allowed16 | {10, 11}
    for tx in range(19):
        for t in range(16):
            used = oracle.intra_tx_type_used(tx, t)
            r = oracle.intra_txt_rate(tx, 0, t)
            if tx in (3, 4, 11, 12, 15, 16, 17, 18) or (tx in (9, 10)):
                assert used == (t == 0) and r == 0
            elif tx in (2,):
                assert used == (t in allowed16) and (r > 0) == used
            else:
                assert used == (t in allowed8) and (r > 0) == used, (tx, t)


def test_intra_mode_rates_match_captures(oracle):
    """The key-frame candidate rates from the default CDFs equal every
    origin block's captured MDS0 luma rate `flr` (no coded neighbour: both
    mode contexts 0), and the reference's pinned entries (leaf_funnel/
    mod.rs:7924-7960: kf_y DC 547 / SMOOTH 1556, V / H with the zero angle
    delta 2874 / 2555, filter-intra flag off 281 and on + FILTER_DC 1803 at
    32x32)."""
    import ctypes
    L = oracle.lib()
    L.or_intra_mode_rate.argtypes = [ctypes.c_int] * 7
    mds0 = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'funnel_mds0.json')))
    n = 0
    for f in mds0['files']:
        for b in f['blocks']:
            if b['org'] != [0, 0]:
                continue
            s = b['sq']
            for (buf, mode, fi, dist, flr, fcr, fc) in b['mds0']:
                assert L.or_intra_mode_rate(s, s, 0, 0, mode, 0, fi) == flr, (f['name'], mode, fi)
                n += 1
    assert n >= 20
    assert L.or_intra_mode_rate(64, 64, 0, 0, 0, 0, 5) == 547
    assert L.or_intra_mode_rate(64, 64, 0, 0, 9, 0, 5) == 1556
    assert L.or_intra_mode_rate(32, 32, 0, 0, 0, 0, 5) - 547 == 281
    assert L.or_intra_mode_rate(32, 32, 0, 0, 0, 0, 0) - 547 == 1803


def test_intra_mode_rates_product_matches_oracle(zav1, oracle):
    import ctypes
    L = oracle.lib()
    L.or_intra_mode_rate.argtypes = [ctypes.c_int] * 7
    P = zav1.lib()
    for (w, h) in ((4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (8, 32), (16, 4)):
        for a in range(5):
            for lc in range(5):
                for m in range(13):
                    for d in ((-3, -1, 0, 2, 3) if 1 <= m <= 8 else (0,)):
                        for fi in ((5, 0, 3) if (m == 0 and max(w, h) <= 32) else (5,)):
                            assert P.zav1_intra_mode_rate(w, h, a, lc, m, d, fi) == \
                                L.or_intra_mode_rate(w, h, a, lc, m, d, fi), (w, h, a, lc, m, d, fi)
    assert P.zav1_intra_mode_rate(64, 64, 0, 0, 0, 0, 0) == zav1.ZAV1_EINVAL
