"""Pin the oracle's scan orders against the reference's C-extracted tables
(digests) and check the fused transform unit against its parts (CPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_scan_orders_match_reference_digests(oracle):
    d = json.load(open(os.path.join(HERE, 'golden', 'scan_golden.json')))
    assert len(d['tables']) == 57
    for t in d['tables']:
        s = oracle.scan(t['tx_size'], t['scan_class']).astype('<u2')
        assert s.size == t['len']
        assert s[:16].tolist() == t['head']
        assert hashlib.sha256(s.tobytes()).hexdigest() == t['sha256'], (t['tx_size'], t['scan_class'])
    lib = oracle.lib()
    assert [lib.or_scan_class(i) for i in range(16)] == d['tx_type_to_scan_index']


def test_product_scan_orders_equal_oracle(zav1, oracle):
    """Host-side (no device) scan generator of the product library."""
    for ts in range(19):
        for c in range(3):
            assert np.array_equal(zav1.scan_order(ts, c), oracle.scan(ts, c))


@pytest.mark.parametrize('tx_size', [0, 2, 4, 9, 12, 17])
def test_tx_unit_composes_parts(oracle, tx_size):
    """dist/eob/level-sum of the fused unit equal the composition of the
    separately pinned fwd, quantize, scan and SSE functions."""
    w, h = oracle.tx_dims(tx_size)
    rng = np.random.default_rng(tx_size)
    res = rng.integers(-300, 300, (h, w)).astype(np.int32)
    for bd, qi in ((8, 80), (10, 160)):
        d, e, s = oracle.tx_unit(res, tx_size, 0, qi, bd)
        c = oracle.fwd_txfm2d(res, tx_size, 0)
        pw, ph = min(w, 32), min(h, 32)
        packed = c[:ph, :pw].copy()
        tq = int(np.sum(c.astype(np.int64) ** 2) - np.sum(packed.astype(np.int64) ** 2))
        ls = [0, 0, 0, 1, 2, 0, 0, 0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 1, 1][tx_size]
        q, dq = oracle.quantize('fp', packed, oracle.quant_row(qi, bd), ls, bd)
        sc = oracle.scan(tx_size, 0)
        nz = np.nonzero(q[sc])[0]
        assert e == (int(nz[-1]) + 1 if nz.size else 0)
        ref = int(np.sum((packed.ravel().astype(np.int64) - dq) ** 2)) + tq
        shift = (1 - ls) * 2
        ref = ref >> shift if shift >= 0 else ref << (-shift)
        assert d == ref
        assert s == int(np.sum(np.abs(q)))


@pytest.mark.parametrize('tx_size', [0, 2, 4, 7, 12, 13])
def test_tx_unit_full_composes_and_reconstructs(oracle, tx_size):
    """The full unit's frequency path equals the residual-level unit (pinned
    above) on src - pred; cul is min(sum |level|, 63) with the DC sign in
    bits 6-7; with no levels the reconstruction is the prediction; the
    spatial distortion is additive over a split crop and vanishes on an
    empty one; at the finest quantizer the reconstruction tracks the source."""
    rng = np.random.default_rng(50 + tx_size)
    w, h = oracle.tx_dims(tx_size)
    for bd in (8, 10):
        mx = (1 << bd) - 1
        dt = np.uint8 if bd == 8 else np.uint16
        src = rng.integers(0, mx + 1, (h, w)).astype(dt)
        pred = np.clip(src.astype(int) + rng.integers(-(6 << (bd - 8)), (6 << (bd - 8)) + 1, (h, w)), 0, mx)
        pred = pred.astype(np.uint16)
        for use_b in (False, True):
            d, e, cul, q, _ = oracle.tx_unit_full(src, pred, tx_size, 0, 120, bd, use_b=use_b, need_recon=False)
            rd, re, rl = oracle.tx_unit(src.astype(np.int32) - pred.astype(np.int32), tx_size, 0, 120, bd,
                                        use_b=use_b)
            assert (d, e) == (rd, re) and rl == int(np.abs(q).sum())
            assert cul & 63 == min(int(np.abs(q).sum()), 63)
            dc = int(q.ravel()[0])
            assert cul >> 6 == (1 if dc < 0 else (2 if dc > 0 else 0))
        # no levels: recon == pred (a flat residual at the coarsest quantizer)
        flat = np.clip(pred.astype(int) + 1, 0, mx).astype(dt)
        d, e, cul, q, rec = oracle.tx_unit_full(flat, pred, tx_size, 0, 255, bd, spatial=True)
        if e == 0:
            assert np.array_equal(rec, pred) and cul == 0
        # spatial distortion: additive over a split crop, zero on an empty crop
        full = oracle.tx_unit_full(src, pred, tx_size, 0, 60, bd, spatial=True)[0]
        top = oracle.tx_unit_full(src, pred, tx_size, 0, 60, bd, spatial=True, crop=(w, h // 2))[0]
        rec = oracle.tx_unit_full(src, pred, tx_size, 0, 60, bd, spatial=True)[4]
        bot = int(((src[h // 2:].astype(np.int64) - rec[h // 2:]) ** 2).sum()) << 4
        assert full == top + bot
        assert oracle.tx_unit_full(src, pred, tx_size, 0, 60, bd, spatial=True, crop=(0, 0))[0] == 0
        # finest quantizer (and no 64-point fold): reconstruction error far
        # below the residual energy
        if w <= 32 and h <= 32:
            d0 = oracle.tx_unit_full(src, pred, tx_size, 0, 0, bd, spatial=True)[0]
            res_e = int(((src.astype(np.int64) - pred) ** 2).sum()) << 4
            assert d0 * 4 < res_e or res_e == 0, (d0, res_e)
