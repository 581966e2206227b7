"""The oracle's RDOQ trellis (optimize_b, enc/quant.rs:1312) pinned by the
reference's own invariant test, a closed-form eob cost, a hand-priced
single-coefficient block and structural invariants over every size and
class.  The multi-coefficient context chain has no captures in the
reference; it is pinned by these invariants and compared GPU == oracle."""
import numpy as np
import pytest

import rdoq_cases as RC
import rdoq_tables as RT


def rdcost(rdmult, rate, dist):
    return ((rate * rdmult + 256) >> 9) + (dist << 7)


def test_reference_dequant_mirror_case(oracle):
    """enc/quant.rs:1852 optimize_keeps_dequant_mirror: the same synthetic
    16x16 spectrum (its LCG), qindex 160, rdoq level 3 (cut-off 60/100),
    rdmult of lambda 248207 (1054880, the reference's rdmult test value);
    the dequant mirror and a nonzero last coefficient hold."""
    n = 256
    t = np.zeros(n, np.int64)
    s = 0x12345678
    for i in range(n):
        s = (s * 1103515245 + 12345) & 0xffffffff
        mag = (s >> 20) % 900
        v = mag if s & 1 == 0 else -mag
        t[i] = int(v / (1 + i // 16))  # Rust integer division truncates
    t = t.astype(np.int32)
    row = oracle.quant_row(160, 8)
    scan = oracle.scan(2, 0)
    q, dq = oracle.quantize('fp', t, row, 0, 8)
    nz = np.nonzero(q[scan])[0]
    eob = int(nz[-1]) + 1
    costs = RT.coeff_costs(160)
    tc, ec = RT.tables_for(costs, 2, 4, 0)
    q2, d2, e2 = oracle.optimize_b(t, q, dq, eob, scan, [row.dequant[0], row.dequant[1]], tc, ec, 1054880, 2, 0,
                                   0, 0, 0, 60, 100)
    deq = np.array([row.dequant[0]] + [row.dequant[1]] * (n - 1), np.int64)
    assert np.array_equal(np.abs(d2), (np.abs(q2).astype(np.int64) * deq).astype(np.int32))
    assert np.array_equal(d2 < 0, q2 < 0)
    assert e2 > 0 and q2[scan[e2 - 1]] != 0
    assert (q2 != q).any()  # the trellis did lower something


def eob_cost_closed_form(eob, tc, ec, tx_class):
    """get_eob_cost from the AV1 eob coding: eob_pt = eob for eob <= 2, else
    ceil(log2(eob)) + 1; the first extra bit with its context cost, the
    other eob_pt - 3 bits as literals."""
    pt = eob if eob <= 2 else int(np.ceil(np.log2(eob))) + 1
    cost = int(ec[11 * (tx_class != 0) + pt - 1])
    if pt > 2:
        start = (1 << (pt - 2)) + 1
        extra = eob - start
        cnt = pt - 3
        cost += int(tc[RT.EOB_EXTRA + 2 * cnt + ((extra >> cnt) & 1)]) + 512 * cnt
    return cost


def test_eob_cost_closed_form(oracle):
    costs = RT.coeff_costs(3)
    tc, ec = RT.tables_for(costs, 3, 6, 1)
    for cls in (0, 1):
        for eob in range(1, 1025):
            assert oracle.rdoq_eob_cost(eob, tc, ec, cls) == eob_cost_closed_form(eob, tc, ec, cls), eob


def golomb(a):
    if a < 15:
        return 0
    r = a - 14
    return 512 * (2 * r.bit_length() - 1)


@pytest.mark.parametrize('tx_size', [0, 2, 3, 4, 5])
def test_single_coefficient_block_priced_by_hand(oracle, tx_size):
    """A block whose only nonzero level is DC (eob 1): keep or lower by one
    by rdcost of (base_eob + dc sign + base range + Golomb, squared error),
    then skip the block when the skip flag beats eob + coefficient + the
    non-skip flag."""
    rng = np.random.default_rng(tx_size)
    w, h = RC.coded_dims(tx_size)
    shift = RC.TX_SCALE[tx_size]
    costs = RT.coeff_costs(40 + tx_size)
    _, txs, ems = oracle.rdoq_size_ctx(tx_size)
    scan = oracle.scan(tx_size, 0)
    seen = set()
    for it in range(300):
        plane = it & 1
        tc, ec = RT.tables_for(costs, txs, ems, plane)
        dqv = [int(rng.integers(8, 1300)), int(rng.integers(8, 1300))]
        a = int(rng.choice([1, 2, 3, 4, 7, 14, 15, 16, 40]))
        sign = int(rng.integers(0, 2))
        t0 = (a * dqv[0] >> shift) + int(rng.integers(-(dqv[0] >> (shift + 1)), (dqv[0] >> (shift + 1)) + 1))
        t0 = -t0 if sign else t0
        t = np.zeros(w * h, np.int32)
        t[0] = t0
        q = np.zeros(w * h, np.int32)
        dq = np.zeros(w * h, np.int32)
        q[0] = -a if sign else a
        dq[0] = -((a * dqv[0]) >> shift) if sign else (a * dqv[0]) >> shift
        rdmult = int(rng.integers(0, 4_000_000))
        skip_ctx, dc_ctx = int(rng.integers(0, 13)), int(rng.integers(0, 3))

        def rate(k):  # the eob coefficient at level k >= 1
            r = int(tc[RT.BASE_EOB + min(k, 3) - 1]) + int(tc[RT.DC_SIGN + 2 * dc_ctx + sign])
            if k > 2:
                r += int(tc[RT.LPS + min(k - 3, 12)]) + golomb(k)
            return r

        def dist(k):
            d = (int(t0) - (-1 if sign else 1) * ((k * dqv[0]) >> shift)) << shift
            return d * d

        d0 = (int(t0) << shift) ** 2
        if a >= 2 and rdcost(rdmult, rate(a - 1), dist(a - 1)) < rdcost(rdmult, rate(a), dist(a)):
            k = a - 1
        else:
            k = a
        acc_rate = eob_cost_closed_form(1, tc, ec, 0) + rate(k)
        skip = rdcost(rdmult, int(tc[RT.SKIP + 2 * skip_ctx + 1]), 0) < \
            rdcost(rdmult, acc_rate + int(tc[RT.SKIP + 2 * skip_ctx]), dist(k) - d0)
        want_q = 0 if skip else (-k if sign else k)
        q2, d2, e2 = oracle.optimize_b(t, q, dq, 1, scan, dqv, tc, ec, rdmult, tx_size, 0, skip_ctx, dc_ctx)
        assert (e2, int(q2[0])) == (0 if skip else 1, want_q), (it, a, rdmult)
        assert int(d2[0]) == (0 if skip else (-1 if sign else 1) * ((k * dqv[0]) >> shift))
        assert not q2[1:].any()
        seen.add((skip, k != a))
    assert len(seen) >= 3  # keep, lower and skip all exercised


@pytest.mark.parametrize('tx_size', list(range(19)))
def test_trellis_invariants(oracle, tx_size):
    """Over every size and class: levels only drop (by one, or to zero past
    the new eob / on skip), signs kept, dequant mirrors the level, the eob
    never grows and its coefficient is nonzero; with rdmult 0 the squared
    error never grows."""
    w, h = RC.coded_dims(tx_size)
    shift = RC.TX_SCALE[tx_size]
    costs = RT.coeff_costs(7 + tx_size)
    _, txs, ems = oracle.rdoq_size_ctx(tx_size)
    for cls in (0, 1, 2):
        for bd, qi in ((8, 60), (10, 200)):
            t, q, dq, eob, scan, dqv = RC.blocks(oracle, tx_size, cls, 6, tx_size * 31 + cls + bd, qi, bd)
            deq = np.array([dqv[0]] + [dqv[1]] * (w * h - 1), np.int64)
            for b in range(len(t)):
                if eob[b] == 0:
                    continue
                for rdmult, cut in ((1054880, (0, 0)), (0, (0, 0)), (6493388, (60, 100))):
                    tc, ec = RT.tables_for(costs, txs, ems, b & 1)
                    q2, d2, e2 = oracle.optimize_b(t[b], q[b], dq[b], eob[b], scan, dqv, tc, ec, rdmult, tx_size,
                                                   cls, b % 13, b % 3, 0, *cut)
                    ch = q2 != q[b]
                    assert np.all((q2[ch] == 0) | (np.abs(q2[ch]) == np.abs(q[b][ch]) - 1))
                    assert np.all((q2[ch] == 0) | (np.sign(q2[ch]) == np.sign(q[b][ch])))
                    mirror = (np.abs(q2).astype(np.int64) * deq) >> shift
                    assert np.array_equal(np.abs(d2[ch]).astype(np.int64), mirror[ch])
                    assert np.array_equal(d2[~ch], dq[b][~ch])
                    assert 0 <= e2 <= eob[b]
                    assert not q2[scan[e2:]].any()
                    assert e2 == 0 or q2[scan[e2 - 1]] != 0
                    if rdmult == 0:
                        err = lambda d: int((((t[b].astype(np.int64) - d) << shift) ** 2).sum())
                        assert err(d2) <= err(dq[b])


@pytest.mark.parametrize('tx_size', [0, 2, 3, 5, 9, 13, 16, 18])
def test_coeff_rate_with_empty_tables_is_literal_bits(oracle, tx_size):
    """With every table entry 0 the coefficient rate is its literal part:
    one bit per nonzero AC sign, Golomb bits of levels >= 15, the eob's
    eob_pt - 3 raw offset bits, and the luma tx-type rate."""
    w, h = RC.coded_dims(tx_size)
    zt, ze = np.zeros(RT.TXB_WORDS, np.int32), np.zeros(RT.EOB_WORDS, np.int32)
    for cls in (0, 1, 2):
        t, q, dq, eob, scan, _ = RC.blocks(oracle, tx_size, cls, 8, 60 + tx_size + cls, 30, 8)
        for b in range(len(q)):
            if eob[b] == 0:
                continue
            lv = np.abs(q[b][scan[:eob[b]]]).astype(np.int64)
            ac_nz = int((lv[1:] != 0).sum())
            gol = sum(golomb(int(a)) for a in lv)
            pt = eob[b] if eob[b] <= 2 else int(np.ceil(np.log2(eob[b]))) + 1
            want = 512 * ac_nz + gol + 512 * max(0, pt - 3) + 777 * (b % 2 == 0)
            got = oracle.cost_coeffs_txb(q[b], eob[b], scan, zt, ze, tx_size, cls, b % 2, 0, 0, 777)
            assert got == want, (tx_size, cls, b)


def test_coeff_rate_dc_only_priced_by_hand(oracle):
    rng = np.random.default_rng(5)
    costs = RT.coeff_costs(55)
    for tx_size in (0, 1, 2, 3, 4, 9):
        w, h = RC.coded_dims(tx_size)
        _, txs, ems = oracle.rdoq_size_ctx(tx_size)
        scan = oracle.scan(tx_size, 0)
        for _ in range(50):
            plane = int(rng.integers(0, 2))
            tc, ec = RT.tables_for(costs, txs, ems, plane)
            a = int(rng.choice([1, 2, 3, 9, 15, 16, 50, 300]))
            sign = int(rng.integers(0, 2))
            q = np.zeros(w * h, np.int32)
            q[0] = -a if sign else a
            skip, dc, txt = int(rng.integers(0, 13)), int(rng.integers(0, 3)), int(rng.integers(0, 4000))
            want = int(tc[RT.SKIP + 2 * skip]) + (txt if plane == 0 else 0) + eob_cost_closed_form(1, tc, ec, 0)
            want += int(tc[RT.BASE_EOB + min(a, 3) - 1]) + int(tc[RT.DC_SIGN + 2 * dc + sign])
            if a > 2:
                want += int(tc[RT.LPS + min(a - 3, 12)]) + golomb(a)
            assert oracle.cost_coeffs_txb(q, 1, scan, tc, ec, tx_size, 0, plane, skip, dc, txt) == want
