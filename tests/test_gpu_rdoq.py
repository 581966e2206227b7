"""GPU parity of the RDOQ trellis (zav1_rdoq_batch) against the oracle's
optimize_b: every tx size, the three tx classes, both plane types, 8 and
10 bits, random rdmult / contexts / sharpness, the full trellis and the
level-2 / level-3 cut-offs; levels, dequantized values and eobs bit-exact."""
import numpy as np
import pytest

import rdoq_cases as RC
import rdoq_tables as RT

pytestmark = pytest.mark.gpu


def workspace(z, tx_size, n_jobs, form):
    """The lane form: a workspace and a threshold of 0 (restored by the
    caller); the wave form: no workspace."""
    if form != 'lane':
        return None
    return z.DeviceArray((max(1, z.lib().zav1_rdoq_workspace(tx_size, n_jobs)),), np.uint8)


def run_case(z, oracle, tx_size, bd, qindex, cut, seed, n_per_class=70, form='wave'):
    prev = z.lib().zav1_rdoq_lane_min(tx_size, 0)
    try:
        return _run_case(z, oracle, tx_size, bd, qindex, cut, seed, n_per_class, form)
    finally:
        z.lib().zav1_rdoq_lane_min(tx_size, prev)


def _run_case(z, oracle, tx_size, bd, qindex, cut, seed, n_per_class, form):
    w, h = RC.coded_dims(tx_size)
    n = w * h
    costs = RT.coeff_costs(seed)
    _, txs, ems = oracle.rdoq_size_ctx(tx_size)
    ts, qs, dqs, eobs, jobs_l = [], [], [], [], []
    scans = np.concatenate([oracle.scan(tx_size, c) for c in (0, 1, 2)]).astype(np.uint16)
    rng = np.random.default_rng(seed)
    dqv = None
    for cls in (0, 1, 2):
        t, q, dq, eob, scan, dqv = RC.blocks(oracle, tx_size, cls, n_per_class, seed * 3 + cls, qindex, bd)
        for b in range(n_per_class):
            j = np.zeros(1, z.RDOQ_JOB)[0]
            j['rdmult'] = int(rng.choice([0, 109013, 1054880, 6493388, int(rng.integers(1, 20_000_000))]))
            j['coeff'] = len(ts) * n
            j['tx_class'] = cls
            j['plane_type'] = int(rng.integers(0, 2))
            j['txb_skip_ctx'] = int(rng.integers(0, 13))
            j['dc_sign_ctx'] = int(rng.integers(0, 3))
            j['sharpness'] = int(rng.random() < 0.15)
            jobs_l.append(j)
            ts.append(t[b]), qs.append(q[b]), dqs.append(dq[b]), eobs.append(eob[b])
    jobs = np.array(jobs_l, dtype=z.RDOQ_JOB)
    T, Q, DQ = np.stack(ts), np.stack(qs), np.stack(dqs)
    E = np.array(eobs, np.uint16)
    D = z.DeviceArray
    dt, dq_, ddq, de = D.from_host(T), D.from_host(Q), D.from_host(DQ), D.from_host(E)
    z.rdoq_batch(tx_size, qindex, bd, dt, dq_, ddq, de, D.from_host(jobs), len(jobs), D.from_host(costs),
                 D.from_host(scans), *cut, d_work=workspace(z, tx_size, len(jobs), form))
    gq, gdq, ge = dq_.to_host(), ddq.to_host(), de.to_host()
    changed = 0
    for b, j in enumerate(jobs):
        cls = int(j['tx_class'])
        scan = scans[RC.SCAN_CLASS[cls] * n:(RC.SCAN_CLASS[cls] + 1) * n]
        if E[b] == 0:
            assert ge[b] == 0 and np.array_equal(gq[b], Q[b]) and np.array_equal(gdq[b], DQ[b])
            continue
        tc, ec = RT.tables_for(costs, txs, ems, int(j['plane_type']))
        rq, rdq, re = oracle.optimize_b(T[b], Q[b], DQ[b], int(E[b]), scan, dqv, tc, ec, int(j['rdmult']), tx_size,
                                        cls, int(j['txb_skip_ctx']), int(j['dc_sign_ctx']), int(j['sharpness']), *cut)
        assert int(ge[b]) == re, (tx_size, b, int(ge[b]), re)
        assert np.array_equal(gq[b], rq), (tx_size, b, np.nonzero(gq[b] != rq)[0][:6])
        assert np.array_equal(gdq[b], rdq), (tx_size, b)
        changed += int((rq != Q[b]).any())
    assert changed > len(jobs) // 4  # the trellis did work on most blocks


@pytest.mark.parametrize('form', ['lane', 'wave'])
@pytest.mark.parametrize('tx_size', list(range(19)))
def test_rdoq_batch_bitexact(gpu, oracle, tx_size, form):
    """Both trellis forms (one lane per block, one wave per block) against
    the oracle."""
    run_case(gpu, oracle, tx_size, 8, 100, (0, 0), 500 + tx_size, form=form)


@pytest.mark.parametrize('tx_size,bd,qindex,cut', [(0, 10, 40, (60, 100)), (2, 10, 200, (80, 100)),
                                                   (3, 8, 20, (60, 100)), (9, 10, 120, (0, 0)),
                                                   (16, 8, 255, (80, 100)), (4, 10, 0, (0, 0))])
@pytest.mark.parametrize('form', ['lane', 'wave'])
def test_rdoq_batch_depths_and_cutoffs(gpu, oracle, tx_size, bd, qindex, cut, form):
    run_case(gpu, oracle, tx_size, bd, qindex, cut, 900 + tx_size * 7 + bd, form=form)


@pytest.mark.parametrize('tx_size', list(range(19)))
def test_coeff_rate_batch_bitexact(gpu, oracle, tx_size):
    """zav1_coeff_rate_batch == cost_coeffs_txb for fp-quantized blocks of
    every class and both plane types, before and after the trellis."""
    z = gpu
    w, h = RC.coded_dims(tx_size)
    n = w * h
    costs = RT.coeff_costs(700 + tx_size)
    _, txs, ems = oracle.rdoq_size_ctx(tx_size)
    scans = np.concatenate([oracle.scan(tx_size, c) for c in (0, 1, 2)]).astype(np.uint16)
    rng = np.random.default_rng(tx_size)
    qs, eobs, jobs = [], [], []
    for cls in (0, 1, 2):
        t, q, dq, eob, scan, dqv = RC.blocks(oracle, tx_size, cls, 40, 70 + tx_size * 3 + cls, 60, 10)
        for b in range(len(q)):
            qb, eb = q[b], int(eob[b])
            plane = int(rng.integers(0, 2))
            if b % 2 and eb:  # half the blocks after the trellis
                tc, ec = RT.tables_for(costs, txs, ems, plane)
                qb, _, eb = oracle.optimize_b(t[b], q[b], dq[b], eb, scan, dqv, tc, ec, 1054880, tx_size, cls)
            j = np.zeros(1, z.RATE_JOB)[0]
            j['coeff'], j['txt_rate'], j['tx_class'], j['plane_type'] = len(qs) * n, int(rng.integers(0, 5000)), cls, plane
            j['txb_skip_ctx'], j['dc_sign_ctx'] = int(rng.integers(0, 13)), int(rng.integers(0, 3))
            qs.append(qb), eobs.append(eb), jobs.append(j)
    Q, E, J = np.stack(qs), np.array(eobs, np.uint16), np.array(jobs, dtype=z.RATE_JOB)
    D = z.DeviceArray
    d_rate = D((len(J),), np.int32)
    z.coeff_rate_batch(tx_size, D.from_host(Q), D.from_host(E), D.from_host(J), len(J), D.from_host(costs),
                       D.from_host(scans), d_rate)
    got = d_rate.to_host()
    for b, j in enumerate(J):
        cls = int(j['tx_class'])
        if E[b] == 0:
            assert got[b] == 0
            continue
        tc, ec = RT.tables_for(costs, txs, ems, int(j['plane_type']))
        sc = scans[RC.SCAN_CLASS[cls] * n:(RC.SCAN_CLASS[cls] + 1) * n]
        want = oracle.cost_coeffs_txb(Q[b], int(E[b]), sc, tc, ec, tx_size, cls, int(j['plane_type']),
                                      int(j['txb_skip_ctx']), int(j['dc_sign_ctx']), int(j['txt_rate']))
        assert int(got[b]) == want, (tx_size, b, int(got[b]), want)


@pytest.mark.parametrize('form', ['lane', 'wave'])
def test_rdoq_and_rate_leave_invalid_jobs_alone(gpu, oracle, form):
    """Blocks whose job fields are out of range (tx class, plane type,
    contexts) are left as they are by the trellis and priced 0 by the rate;
    their neighbours in the batch are processed normally."""
    z = gpu
    tx_size, n = 1, 64
    costs = RT.coeff_costs(31)
    _, txs, ems = oracle.rdoq_size_ctx(tx_size)
    scans = np.concatenate([oracle.scan(tx_size, c) for c in (0, 1, 2)]).astype(np.uint16)
    t, q, dq, eob, scan, dqv = RC.blocks(oracle, tx_size, 0, 8, 77, 40, 8)
    jobs = np.zeros(8, z.RDOQ_JOB)
    jobs['rdmult'] = 1054880
    jobs['coeff'] = np.arange(8) * n
    bad = {1: ('tx_class', 3), 3: ('plane_type', 2), 5: ('txb_skip_ctx', 13), 7: ('dc_sign_ctx', 3)}
    for b, (f, v) in bad.items():
        jobs[b][f] = v
    rjobs = np.zeros(8, z.RATE_JOB)
    rjobs['coeff'] = jobs['coeff']
    for b, (f, v) in bad.items():
        rjobs[b][f] = v
    D = z.DeviceArray
    dq_, ddq, de = D.from_host(q), D.from_host(dq), D.from_host(eob.astype(np.uint16))
    dcosts, dscans = D.from_host(costs), D.from_host(scans)
    prev = z.lib().zav1_rdoq_lane_min(tx_size, 0)
    try:
        z.rdoq_batch(tx_size, 40, 8, D.from_host(t), dq_, ddq, de, D.from_host(jobs), 8, dcosts, dscans,
                     d_work=workspace(z, tx_size, 8, form))
    finally:
        z.lib().zav1_rdoq_lane_min(tx_size, prev)
    gq, ge = dq_.to_host(), de.to_host()
    d_rate = D((8,), np.int32)
    z.coeff_rate_batch(tx_size, dq_, de, D.from_host(rjobs), 8, dcosts, dscans, d_rate)
    rate = d_rate.to_host()
    tc, ec = RT.tables_for(costs, txs, ems, 0)
    for b in range(8):
        if b in bad:
            assert np.array_equal(gq[b], q[b]) and ge[b] == eob[b], b
            assert rate[b] == 0, b
        elif eob[b]:
            rq, _, re = oracle.optimize_b(t[b], q[b], dq[b], int(eob[b]), scan, dqv, tc, ec, 1054880, tx_size, 0)
            assert np.array_equal(gq[b], rq) and ge[b] == re, b
            assert rate[b] == (oracle.cost_coeffs_txb(rq, re, scan, tc, ec, tx_size, 0) if re else 0), b
