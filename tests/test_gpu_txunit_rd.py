"""GPU parity of the full-RD transform unit (zav1_txunit_rd_batch: forward ->
quantize_b or quantize_fp + RDOQ -> rate -> frequency or spatial distortion,
recon) against the oracle chain, bit-exact at 8 and 10 bits on every tx
size, with the default-CDF rate tables; and the C encoder's captured
full-loop records at the frame origin reproduced through the device."""
import json
import os

import numpy as np
import pytest

from oracle_chain import rd_unit, tx_class  # noqa: F401

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32), (32, 16),
           (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]
MODES = [(1, None, False), (1 | 2, None, True), (8, (0, 0), False), (8 | 2, (60, 100), True), (8 | 4, (80, 100), False)]


def types_for(w, h):
    if max(w, h) == 64:
        return [0, 9]
    if max(w, h) == 32:
        return [0, 9, 10, 11]
    return list(range(16))


def run_rd(z, d_src, stride, pw, ph, bd, ts, preds, jobs, qidx, flags, cut, d_costs, want_q=True, want_rec=True):
    w, h = TX_DIMS[ts]
    n = len(jobs)
    d_pred = z.DeviceArray.from_host(np.concatenate([p.ravel() for p in preds]).astype(np.uint16))
    d_jobs = z.DeviceArray.from_host(jobs.view(np.uint8))
    d_work = z.DeviceArray(max(z.txunit_rd_workspace(ts, n), 1), np.uint8)
    d_dist, d_eob = z.DeviceArray(n, np.uint64), z.DeviceArray(n, np.int32)
    d_rate, d_cul = z.DeviceArray(n, np.int32), z.DeviceArray(n, np.uint8)
    d_q = z.DeviceArray(n * min(w, 32) * min(h, 32), np.int32) if want_q else None
    d_rec = z.DeviceArray(n * w * h, np.uint16) if want_rec else None
    z.txunit_rd_batch(d_src, stride, pw, ph, bd, ts, d_pred, d_jobs, n, qidx, flags, d_costs, d_work, d_dist, d_eob,
                      d_rate, d_cul, d_q, d_rec, cut_off=cut if cut else (0, 0))
    z.sync()
    out = dict(dist=d_dist.to_host(), eob=d_eob.to_host(), rate=d_rate.to_host(), cul=d_cul.to_host())
    if want_q:
        out['q'] = d_q.to_host().reshape(n, min(h, 32), min(w, 32))
    if want_rec:
        out['rec'] = d_rec.to_host().reshape(n, h, w)
    return out


@pytest.mark.parametrize('bd', [8, 10])
@pytest.mark.parametrize('mode', range(len(MODES)))
def test_txunit_rd_bitexact(gpu, oracle, bd, mode):
    z = gpu
    flags, cut, spatial = MODES[mode]
    rng = np.random.default_rng(bd * 10 + mode)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    PH, PW = 160, 192
    yy, xx = np.mgrid[0:PH, 0:PW]
    plane = (np.sin(xx / 9.0) * np.cos(yy / 13.0) * 0.4 + 0.5) * mx + rng.normal(0, 5 << (bd - 8), (PH, PW))
    plane = np.clip(plane, 0, mx).astype(dt)
    d_src = z.DeviceArray.from_host(plane)
    qidx = int(rng.integers(40, 220))
    d_costs = z.DeviceArray.from_host(z.build_coeff_costs(qidx))
    tables = oracle.build_coeff_costs(qidx)
    lam = int(rng.integers(20000, 400000))
    for ts, (w, h) in enumerate(TX_DIMS):
        n = 10
        jobs = np.zeros(n, z.TXUNIT_RD_JOB)
        preds = []
        types = types_for(w, h)
        for j in range(n):
            x, y = int(rng.integers(0, PW - w + 1)), int(rng.integers(0, PH - h + 1))
            blk = plane[y:y + h, x:x + w].astype(np.int64)
            amp = (10 + 6 * (j % 4)) << (bd - 8)
            p = blk + rng.integers(-amp, amp + 1, blk.shape) if j % 5 else np.full_like(blk, int(blk.mean()))
            preds.append(np.clip(p, 0, mx).astype(np.uint16))
            t = types[j % len(types)]
            plane_t = j % 2 if j < 8 else 0
            jobs[j] = (x, y, j * w * h, t, int(rng.integers(1, w + 1)), int(rng.integers(1, h + 1)), plane_t,
                       int(rng.integers(0, 13)), int(rng.integers(0, 3)), 0, 0,
                       oracle.intra_txt_rate(ts, j % 13, t) if plane_t == 0 else 0,
                       ((lam * (17 if plane_t == 0 else 13) + 2) >> 2))
        if max(w, h) >= 32:
            jobs[3]['x'] = PW - w + 1          # out of bounds: eob -22, rate 0
        got = run_rd(z, d_src, PW, PW, PH, bd, ts, preds, jobs, qidx, flags, cut, d_costs)
        for j in range(n):
            jb = jobs[j]
            if j == 3 and max(w, h) >= 32:
                assert got['eob'][j] == -22 and got['rate'][j] == 0
                continue
            x, y = int(jb['x']), int(jb['y'])
            ref = rd_unit(oracle, plane[y:y + h, x:x + w], preds[j].astype(np.int64), ts, int(jb['tx_type']), qidx,
                          bd, tables, rdoq=cut if flags & 8 else None, rdmult=int(jb['rdmult']),
                          tsc=int(jb['txb_skip_ctx']), dsc=int(jb['dc_sign_ctx']), txt_rate=int(jb['txt_rate']),
                          plane=int(jb['plane_type']), spatial=spatial, crop=(int(jb['crop_w']), int(jb['crop_h'])))
            eob, rate, dist, cul, qa, rec = ref
            ctx = (bd, mode, ts, j, int(jb['tx_type']))
            assert (int(got['eob'][j]), int(got['rate'][j]), int(got['dist'][j]), int(got['cul'][j])) == \
                (eob, rate, dist, cul), ctx
            assert np.array_equal(got['q'][j], qa), ctx
            if flags & (2 | 4):
                assert np.array_equal(got['rec'][j], rec), ctx


def test_txunit_rd_rate_est2(gpu):
    """ZAV1_TXU_RATE_EST2 (coeff_rate_est_lvl 2, tx_pipeline.rs:605): a luma
    unit with eob < (w*h) >> 6 is priced 6000 + 1000 * eob, every other unit
    (and every other output) is the exact run's."""
    z = gpu
    rng = np.random.default_rng(31)
    PH, PW = 128, 128
    plane = np.clip(rng.normal(512, 40, (PH, PW)), 0, 1023).astype(np.uint16)
    d_src = z.DeviceArray.from_host(plane)
    qidx = 180
    d_costs = z.DeviceArray.from_host(z.build_coeff_costs(qidx))
    for ts in (1, 2, 3, 4, 9):
        w, h = TX_DIMS[ts]
        n = 24
        jobs = np.zeros(n, z.TXUNIT_RD_JOB)
        preds = []
        for j in range(n):
            x, y = int(rng.integers(0, PW - w + 1)), int(rng.integers(0, PH - h + 1))
            amp = 4 + 12 * (j % 6)
            preds.append(np.clip(plane[y:y + h, x:x + w].astype(np.int64) + rng.integers(-amp, amp + 1, (h, w)),
                                 0, 1023).astype(np.uint16))
            jobs[j] = (x, y, j * w * h, 0, w, h, j % 3 == 2, 0, 0, 0, 0, 40, 100000)
        for flags in (1, 8 | 2):
            a = run_rd(z, d_src, PW, PW, PH, 10, ts, preds, jobs, qidx, flags, (0, 0), d_costs)
            b = run_rd(z, d_src, PW, PW, PH, 10, ts, preds, jobs, qidx, flags | 16, (0, 0), d_costs)
            small = (a['eob'] < (w * h) >> 6) & (jobs['plane_type'] == 0)
            assert small.any() and (~small).any(), (ts, flags)
            assert np.array_equal(b['rate'][small], 6000 + 1000 * a['eob'][small]), (ts, flags)
            assert np.array_equal(b['rate'][~small], a['rate'][~small]), (ts, flags)
            for k in ('eob', 'dist', 'cul', 'q', 'rec'):
                assert np.array_equal(a[k], b[k]), (ts, flags, k)


# --- the C encoder's records at the frame origin, through the device --------

RDOQ_LEVEL = {('gradient_64x64', 20): 0, ('gradient_64x64', 40): 3, ('gradient_64x64', 55): 3,
              ('gradient_128x128', 20): 2, ('gradient_128x128', 40): 2, ('gradient_128x128', 55): 2}
CUTOFF = {2: (80, 100), 3: (60, 100)}
FIMODE_TO_INTRADIR = [0, 1, 2, 6, 0]


def test_c_capture_records_through_device(gpu):
    z = gpu
    full = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'funnel_full.json')))
    A = z.DeviceArray
    checked = 0
    for f in full['files']:
        r, c = np.mgrid[0:f['h'], 0:f['w']]
        img = (((r * 255) // f['h']).astype(np.uint8) ^ ((c * 3) & 0x3f).astype(np.uint8)).astype(np.uint8)
        d_img = A.from_host(img)
        b = f['blocks'][0]
        qidx, lam = b['qidx'], b['lambda']
        lvl = RDOQ_LEVEL[(f['name'].split('_q')[0], f['qp'])]
        d_costs = A.from_host(z.build_coeff_costs(qidx))

        def predict(side, cands):
            ij = np.zeros(len(cands), z.INTRA_JOB)
            for m, (mode, fi) in enumerate(cands):
                ij[m] = (0, 0, side, side, mode, 0, fi, 1, 0, 0, 0, 0, 0, 0, m * side * side, 0)
            d_pred = A(len(cands) * side * side, np.uint16)
            z.intra_pred_batch(d_img, f['w'], 8, A.from_host(ij.view(np.uint8)), len(cands), d_pred)
            return d_pred.to_host().reshape(len(cands), side, side)

        # MDS1: DCT_DCT quantize_b, frequency distortion, depth 0
        side = b['sq']
        ts = {16: 2, 32: 3, 64: 4}[side]
        cands = [(m['mode'], m['fi']) for m in b['mds1']]
        preds = list(predict(side, cands))
        jobs = np.zeros(len(cands), z.TXUNIT_RD_JOB)
        for j, (mode, fi) in enumerate(cands):
            jobs[j] = (0, 0, j * side * side, 0, side, side, 0, 0, 0, 0, 0, 0, 0)
        got = run_rd(z, d_img, f['w'], f['w'], f['h'], 8, ts, preds, jobs, qidx, 1, None, d_costs, False, False)
        for j, m in enumerate(b['mds1']):
            assert (int(got['eob'][j]), int(got['rate'][j]), int(got['dist'][j])) == \
                (m['eob0'], m['ybits'], m['ydist']), (f['name'], m)
            checked += 1
        # MDS3 tx search records: depth 0 and the depth-1 unit at the origin
        for dep in (0, 1):
            sd = side >> dep
            tsd = {16: 2, 32: 3, 64: 4}[sd]
            recs = []
            for cand in b['txt']:
                for t in cand['records']:
                    if t['kind'] == 'txt' and t['dep'] == dep and t['txb'] == 0 and t['tt'] == 0:
                        recs.append((cand['mode'], cand['fi'], t))
            preds = list(predict(sd, [(m, fi) for m, fi, _ in recs]))
            jobs = np.zeros(len(recs), z.TXUNIT_RD_JOB)
            for j, (mode, fi, t) in enumerate(recs):
                rate_t = z.intra_txt_rate(tsd, FIMODE_TO_INTRADIR[fi] if fi < 5 else mode, 0)
                jobs[j] = (0, 0, j * sd * sd, 0, sd, sd, 0, t['tsc'], t['dsc'], 0, 0, rate_t, (lam * 17 + 2) >> 2)
            flags = (8 if lvl else 1) | 2
            got = run_rd(z, d_img, f['w'], f['w'], f['h'], 8, tsd, preds, jobs, qidx, flags, CUTOFF.get(lvl), d_costs,
                         False, False)
            for j, (mode, fi, t) in enumerate(recs):
                assert (int(got['eob'][j]), int(got['rate'][j]), int(got['dist'][j])) == \
                    (t['eob'], t['bits'], t['dist']), (f['name'], mode, fi, t)
                checked += 1
    want = sum(len(f['blocks'][0]['mds1']) + 2 * len(f['blocks'][0]['txt']) for f in full['files'])
    assert checked == want == 65
