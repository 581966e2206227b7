"""Pin the oracle's transforms, quantizers and SATD against the reference's
known-answer vectors, closed forms and brute force (CPU only)."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _kind(func):
    return 0 if 'dct' in func else (1 if 'adst' in func else 3)


def test_txfm1d_golden_vectors(oracle):
    """40 vectors measured from C SVT-AV1 (golden_parity.rs:36-720)."""
    d = json.load(open(os.path.join(HERE, 'golden', 'txfm1d_golden.json')))
    assert len(d['vectors']) == 40
    for v in d['vectors']:
        f = v['func']
        if f.startswith('f'):
            out = oracle.fwd_txfm1d(_kind(f), v['input'], v['arg'])
        else:
            out = oracle.inv_txfm1d(_kind(f), v['input'], v['arg'])
        assert out is not None, v['test']
        assert list(out) == v['golden'], (v['test'], list(out), v['golden'])


def test_cospi_closed_form_and_c_spot_values(oracle):
    """COSPI_ARR_DATA = round(cos(j*pi/128)*2^(10+i)) (fwd_txfm.rs:15-18);
    C spot values golden_parity.rs:229-250."""
    for cb in range(10, 17):
        row = oracle.lib().or_cospi_row(cb)
        for j in range(64):
            assert row[j] == int(math.floor(math.cos(j * math.pi / 128) * (1 << cb) + 0.5))
    q12 = oracle.lib().or_cospi_row(12)
    assert (q12[0], q12[16], q12[32], q12[48], q12[63]) == (4096, 3784, 2896, 1567, 101)
    s12 = oracle.lib().or_sinpi_row(12)
    assert [s12[i] for i in range(5)] == [0, 1321, 2482, 3344, 3803]


@pytest.mark.parametrize('n', [4, 8, 16, 32, 64])
def test_fdct_matches_float_dct(oracle, n):
    """Integer DCT ~ scaled orthonormal DCT-II: a dropped term, wrong sign or
    wrong output index moves a coefficient far outside the rounding slack."""
    rng = np.random.default_rng(n)
    k = np.arange(n)
    basis = np.cos(np.pi * (2 * k[None, :] + 1) * k[:, None] / (2 * n))
    for _ in range(50):
        x = rng.integers(-255, 256, n)
        out = oracle.fwd_txfm1d(0, x, 13)
        ref = basis @ x * math.sqrt(2.0 / n) * math.sqrt(n / 2.0)  # AV1 scale: DC = sum/sqrt(2)
        ref[0] /= math.sqrt(2.0)
        assert np.max(np.abs(out - ref)) <= 2 + n * 0.05, (n, np.max(np.abs(out - ref)))


@pytest.mark.parametrize('n', [4, 8, 16])
def test_fadst_inverts(oracle, n):
    """iadst(fadst(x)) reproduces x up to the known gain and rounding."""
    rng = np.random.default_rng(100 + n)
    for _ in range(50):
        x = rng.integers(-255, 256, n)
        f = oracle.fwd_txfm1d(1, x, 12)
        y = oracle.inv_txfm1d(1, f, 31)
        gain = n / 2.0
        assert np.max(np.abs(y - x * gain)) <= 1 + n, (n, y, x)


ALL_TX = list(range(19))


def _valid_types(tx_size, oracle):
    w, h = oracle.tx_dims(tx_size)
    out = []
    for t in range(16):
        col = [0, 1, 0, 1, 2, 0, 2, 1, 2, 3, 0, 3, 1, 3, 2, 3][t]
        row = [0, 0, 1, 1, 0, 2, 2, 2, 1, 3, 3, 0, 3, 1, 3, 2][t]
        if (col in (1, 2) and h > 16) or (row in (1, 2) and w > 16):
            continue
        out.append(t)
    return out


@pytest.mark.parametrize('tx_size', ALL_TX)
def test_fwd2d_separable_dc_and_flips(oracle, tx_size):
    """2D composition invariants: constant input puts all energy in DC for
    DCT_DCT; FLIPADST_DCT of x equals ADST_DCT of x flipped upside down."""
    w, h = oracle.tx_dims(tx_size)
    blk = np.full((h, w), 37, dtype=np.int32)
    out = oracle.fwd_txfm2d(blk, tx_size, 0)
    assert out[0, 0] != 0 and np.count_nonzero(out) == 1
    if h <= 16:
        rng = np.random.default_rng(tx_size)
        x = rng.integers(-255, 256, (h, w)).astype(np.int32)
        a = oracle.fwd_txfm2d(x[::-1].copy(), tx_size, 1)   # ADST_DCT on flipped rows
        b = oracle.fwd_txfm2d(x, tx_size, 4)                 # FLIPADST_DCT
        assert np.array_equal(a, b)
    if w <= 16:
        rng = np.random.default_rng(50 + tx_size)
        x = rng.integers(-255, 256, (h, w)).astype(np.int32)
        a = oracle.fwd_txfm2d(x[:, ::-1].copy(), tx_size, 2)  # DCT_ADST on flipped columns
        b = oracle.fwd_txfm2d(x, tx_size, 5)                  # DCT_FLIPADST
        assert np.array_equal(a, b)


@pytest.mark.parametrize('tx_size', ALL_TX)
def test_inv2d_roundtrip(oracle, tx_size):
    """inverse(forward(x)) ~ x for every legal type (gain removed by the shifts)."""
    w, h = oracle.tx_dims(tx_size)
    rng = np.random.default_rng(7 + tx_size)
    # AV1 codes 64-point dimensions with DCT only (the inverse reads 32x32)
    types = [0] if max(w, h) == 64 else _valid_types(tx_size, oracle)
    for t in types:
        x = rng.integers(-60, 61, (h, w)).astype(np.int32)
        if w == 64 or h == 64:
            # 64-point inverses read only the low 32x32 band: use band-limited input
            k1, k2 = rng.integers(0, 8, 2)
            r = np.arange(h)[:, None]
            c = np.arange(w)[None, :]
            x = np.round(50 * np.cos(np.pi * (2 * r + 1) * k1 / (2 * h)) *
                         np.cos(np.pi * (2 * c + 1) * k2 / (2 * w))).astype(np.int32)
        c = oracle.fwd_txfm2d(x, tx_size, t)
        assert c is not None
        y = oracle.inv_txfm2d(c, tx_size, t, 8)
        tol = 3 if max(w, h) < 64 else 12
        assert np.max(np.abs(y - x)) <= tol, (tx_size, t, np.max(np.abs(y - x)))


def test_unsupported_combinations(oracle):
    assert oracle.fwd_txfm2d(np.zeros((32, 32), np.int32), 3, 1) is None   # ADST on 32
    assert oracle.inv_txfm2d(np.zeros((64, 64), np.int32), 4, 1, 8) is None  # ADST on 64
    assert oracle.inv_txfm2d(np.zeros((64, 64), np.int32), 4, 9, 8) is not None  # identity 64 exists


def test_quant_row_golden(oracle):
    d = json.load(open(os.path.join(HERE, 'golden', 'quant_row_golden.json')))
    for g in d['rows']:
        r = oracle.quant_row(g['qindex'], g['bd'])
        for f in ('zbin', 'round', 'quant', 'quant_shift', 'quant_fp', 'round_fp', 'dequant'):
            assert list(getattr(r, f)) == g[f], f


def test_quant_row_tables(oracle):
    q = json.load(open(os.path.join(HERE, 'golden', 'av1_qlookup.json')))
    for qi in (0, 1, 20, 40, 55, 128, 200, 255):
        r = oracle.quant_row(qi, 10)
        assert list(r.dequant) == [q['dc10'][qi], q['ac10'][qi]]
        r8 = oracle.quant_row(qi, 8)
        assert list(r8.dequant) == [q['dc8'][qi], q['ac8'][qi]]


def _brute_fp(c, row, ls, bd, iz):
    """Independent integer re-derivation of the fp quantizer for one value by
    exhaustive search over levels (smallest level whose dequant window holds
    the scaled magnitude), used on values where both must agree."""
    a = abs(c)
    if (a << (1 + ls)) < row.dequant[iz]:
        return 0
    rnd = (row.round_fp[iz] + ((1 << ls) >> 1)) >> ls
    x = a + rnd
    if bd == 8:
        x = min(x, 32767)
    return (x * row.quant_fp[iz]) >> (16 - ls)


def test_quantize_properties(oracle):
    """Dequant mirror, sign symmetry, dead zone, monotone levels, bd8 clamp vs
    bd10 divergence above INT16 (quant.rs:1766-1799)."""
    rng = np.random.default_rng(3)
    for bd in (8, 10):
        for qi in (20, 100, 160, 220):
            row = oracle.quant_row(qi, bd)
            for ls in (0, 1, 2):
                c = rng.integers(-40000 if bd == 10 else -20000, 40000 if bd == 10 else 20000, 1024).astype(np.int32)
                c[0] = abs(c[0])
                for kind in ('fp', 'b'):
                    q, dq = oracle.quantize(kind, c, row, ls, bd)
                    iz = (np.arange(c.size) != 0).astype(int)
                    deq = np.array(row.dequant)[iz]
                    assert np.array_equal(np.abs(dq), (np.abs(q).astype(np.int64) * deq) >> ls)
                    assert np.all((q == 0) | (np.sign(q) == np.sign(c)))
                    qn, _ = oracle.quantize(kind, -c, row, ls, bd)
                    assert np.array_equal(qn, -q)
                    order = np.argsort(np.abs(c[1:]))
                    assert np.all(np.diff(np.abs(q[1:])[order]) >= 0)
                q, _ = oracle.quantize('fp', c, row, ls, bd)
                exp = np.array([_brute_fp(int(v), row, ls, bd, int(i != 0)) for i, v in enumerate(c)])
                assert np.array_equal(np.abs(q), exp)
    row = oracle.quant_row(160, 10)
    big = np.zeros(16, np.int32)
    big[0] = 90000
    q8, _ = oracle.quantize('fp', big, row, 1, 8)
    q10, _ = oracle.quantize('fp', big, row, 1, 10)
    assert abs(q10[0]) > abs(q8[0])


def test_sse_wrapping(oracle):
    a = np.array([2 ** 31 - 1, -(2 ** 31)], dtype=np.int32)
    b = np.array([-(2 ** 31), 2 ** 31 - 1], dtype=np.int32)
    e = (2 ** 32 - 1)
    assert oracle.sse_i32(a, b) == (2 * e * e) % (1 << 64)
    x = np.arange(-100, 100, dtype=np.int32)
    assert oracle.sq_sum_i32(x) == int(np.sum(x.astype(np.int64) ** 2))


def _walsh(n):
    h = np.array([[1]])
    while h.shape[0] < n:
        h = np.block([[h, h], [h, -h]])
    return h


@pytest.mark.parametrize('n', [4, 8, 16, 32])
def test_hadamard_matches_walsh(oracle, n):
    """Small-range input (no wrap): the coefficient multiset and SATD equal a
    brute-force Walsh-Hadamard transform with the per-size output scaling
    (4x4 /4, 16x16 /2, 32x32 /8, as the halving stages imply)."""
    rng = np.random.default_rng(n)
    for _ in range(20):
        x = rng.integers(-40, 41, (n, n)).astype(np.int16)
        c = oracle.hadamard(n, x)
        H = _walsh(n)
        full = H @ x.astype(np.int64) @ H.T
        scale = {4: 4, 8: 1, 16: 2, 32: 8}[n]
        ref = np.sort(np.abs(full).ravel())
        got = np.sort(np.abs(c.astype(np.int64)).ravel() * scale)
        # the >>1 / >>2 stages truncate; each coefficient lies within the slack
        assert np.max(np.abs(got - ref)) <= 4 * scale, n
        if n == 8:
            assert np.array_equal(got, ref)


_WORST_L1 = {}


def _coefficient_bound(oracle, amp):
    """Largest |coefficient| any size / type can produce from residuals in
    [-amp, amp]: the forward transform is linear up to per-stage rounding, so
    it is the l1 norm of a transform row times amp (measured from impulse
    responses of amplitude 1023, scaled), plus a margin for the roundings."""
    if 'l1' not in _WORST_L1:
        worst, margin = 0.0, 0
        for ts in range(19):
            w, h = oracle.tx_dims(ts)
            for t in range(16):
                if oracle.fwd_txfm2d(np.zeros((h, w), np.int32), ts, t) is None:
                    continue
                M = np.zeros((h * w, h * w))
                for i in range(h * w):
                    e = np.zeros(h * w, np.int32)
                    e[i] = 1023
                    M[:, i] = oracle.fwd_txfm2d(e.reshape(h, w), ts, t).ravel()
                worst = max(worst, np.abs(M).sum(axis=1).max() / 1023.0)
                margin = max(margin, 64 * h * w)
        _WORST_L1['l1'] = (worst, margin)
    worst, margin = _WORST_L1['l1']
    return int(np.ceil(worst * amp)) + margin


def test_fused_kernel_coefficient_bound(oracle):
    """The fused search / unit kernels quantize in 32-bit arithmetic under the
    contract |coefficient| < 2^21 for residuals of bd <= 10 planes: the
    bound must stay far below 2^21 for every size and type."""
    worst = _coefficient_bound(oracle, 1023)
    assert worst < (1 << 19), worst


def test_fused_fp_quantizer_error_and_level_bounds(oracle):
    """The fused kernels' 32-bit fp quantizer relies on two facts about every
    quantizer row (bd 8 and 10, all 256 qindex): quant_fp <= 2^14, so the
    level (x * quant_fp) >> (16 - ls) is the high word of
    (x << ls) * (quant_fp << 16) with both factors below 2^32; and the
    quantization error |coef| - |dq| of every reachable magnitude stays below
    2^15, so its square is a 24-bit-operand product that fits 32 bits.
    The error bound is checked on the oracle quantizer (DC and AC parameters,
    every log_scale) over every magnitude a residual of the bit depth can
    produce: l1 * (2^bd - 1) plus a rounding allowance (the largest row l1
    norm is 128.13; the sign-aligned worst residual measures 32648 at 8 bits
    against the linear 32672, so rounding moves a coefficient by far less
    than the 64 / 256 allowed), i.e. below the 8-bit INT16 clamp and below
    2^17 + 2^12 at 10 bits."""
    _coefficient_bound(oracle, 1)
    l1 = _WORST_L1['l1'][0]
    reach = {8: 32767, 10: (1 << 17) + (1 << 12)}
    assert l1 * 255 + 64 <= reach[8] and l1 * 1023 + 256 <= reach[10], l1
    for bd in (8, 10):
        a = np.arange(0, reach[bd] + 1, dtype=np.int32)
        for qi in range(256):
            row = oracle.quant_row(qi, bd)
            assert max(row.quant_fp[0], row.quant_fp[1]) <= (1 << 14), (bd, qi)
            for which in (0, 1):
                r = oracle.QuantRow()
                for f, _ in oracle.QuantRow._fields_:
                    v = getattr(row, f)[which]
                    getattr(r, f)[0] = v
                    getattr(r, f)[1] = v
                dqv = row.dequant[which]
                assert dqv < (1 << 13)
                for ls in (0, 1, 2):
                    q, dq = oracle.quantize('fp', a, r, ls, bd)
                    e = a.astype(np.int64) - np.abs(dq).astype(np.int64)
                    # one step, plus the truncation of quant_fp = 2^16 / dequant (a * dequant / 2^16)
                    # and the rounding offset where the 8-bit clamp bites
                    lim = dqv + row.round_fp[which] + (int(a[-1]) * dqv >> 16) + 1
                    assert e.max() <= lim and e.min() >= -lim, (bd, qi, which, ls, e.min(), e.max())
                    assert np.abs(e).max() < (1 << 15)
