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
