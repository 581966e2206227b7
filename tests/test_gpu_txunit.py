"""GPU parity of the fused transform unit (residual -> forward -> fold ->
quantize fp/b -> eob / cul / frequency or spatial distortion, inverse ->
recon) against the oracle, bit-exact, every tx size at 8 and 10 bits, with
illegal types and out-of-bounds units reported per job."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32), (32, 16),
           (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def legal(w, h):
    if w == 64 or h == 64:
        return [0, 9]
    if w == 32 or h == 32:
        return [0, 9, 10, 11]
    return list(range(16))


@pytest.mark.parametrize('bd', [8, 10])
@pytest.mark.parametrize('flags', [0, 1, 2 | 4, 1 | 2, 4])
def test_txunit_batch_bitexact(gpu, oracle, bd, flags):
    z = gpu
    rng = np.random.default_rng(bd * 100 + flags)
    mx = (1 << bd) - 1
    dt = np.uint8 if bd == 8 else np.uint16
    PH, PW = 192, 224
    yy, xx = np.mgrid[0:PH, 0:PW]
    plane = (np.sin(xx / 7.0) * np.cos(yy / 11.0) * 0.4 + 0.5) * mx + rng.normal(0, 6 << (bd - 8), (PH, PW))
    plane = np.clip(plane, 0, mx).astype(dt)
    d_src = z.DeviceArray.from_host(plane)
    qindex = int(rng.integers(20, 200))
    for ts, (w, h) in enumerate(TX_DIMS):
        n = 24
        jobs = np.zeros(n, z.TXUNIT_JOB)
        preds = []
        types = legal(w, h)
        for j in range(n):
            x, y = int(rng.integers(0, PW - w + 1)), int(rng.integers(0, PH - h + 1))
            blk = plane[y:y + h, x:x + w].astype(np.int64)
            # a prediction that is close (smooth) to the block, occasionally exact
            p = blk + rng.integers(-(12 << (bd - 8)), (12 << (bd - 8)) + 1, blk.shape) if j % 7 else blk.copy()
            preds.append(np.clip(p, 0, mx).astype(np.uint16))
            t = types[j % len(types)]
            jobs[j] = (x, y, j * w * h, t, int(rng.integers(1, w + 1)), int(rng.integers(1, h + 1)), 0)
        # an illegal type (ADST on a 32/64-point dimension) and an out-of-bounds unit
        if h > 16:
            jobs[1]['tx_type'] = 1   # ADST columns
        elif w > 16:
            jobs[1]['tx_type'] = 2   # ADST rows
        jobs[2]['x'] = PW - w + 1
        pred = np.concatenate([p.ravel() for p in preds])
        d_pred = z.DeviceArray.from_host(pred)
        d_jobs = z.DeviceArray.from_host(jobs.view(np.uint8))
        d_dist, d_eob, d_cul = z.DeviceArray(n, np.uint64), z.DeviceArray(n, np.int32), z.DeviceArray(n, np.uint8)
        pw_, ph_ = min(w, 32), min(h, 32)
        d_q = z.DeviceArray(n * pw_ * ph_, np.int32)
        d_rec = z.DeviceArray(n * w * h, np.uint16)
        d_q.zero()
        d_rec.zero()
        z.txunit_batch(d_src, PW, PW, PH, bd, ts, d_pred, d_jobs, n, qindex, flags, d_dist, d_eob, d_cul, d_q,
                       d_rec)
        gd, ge, gc = d_dist.to_host(), d_eob.to_host(), d_cul.to_host()
        gq = d_q.to_host().reshape(n, ph_, pw_)
        gr = d_rec.to_host().reshape(n, h, w)
        assert ge[2] == -22 and gd[2] == 0
        if w > 16 or h > 16:
            assert ge[1] == -1 and gd[1] == 0
        for j in range(n):
            if ge[j] < 0:
                continue
            x, y = int(jobs['x'][j]), int(jobs['y'][j])
            blk = np.ascontiguousarray(plane[y:y + h, x:x + w])
            recon = bool(flags & 6)
            ref = oracle.tx_unit_full(blk, preds[j], ts, int(jobs['tx_type'][j]), qindex, bd, use_b=bool(flags & 1),
                                      need_recon=recon, spatial=bool(flags & 2),
                                      crop=(int(jobs['crop_w'][j]), int(jobs['crop_h'][j])))
            rd, re, rc, rq, rr = ref
            tag = (ts, j, int(jobs['tx_type'][j]), flags, bd)
            assert (int(gd[j]), int(ge[j]), int(gc[j])) == (rd, re, rc), (tag, gd[j], ge[j], gc[j], rd, re, rc)
            assert np.array_equal(gq[j], rq), tag
            if recon:
                assert np.array_equal(gr[j], rr), (tag, np.argwhere(gr[j] != rr)[:4].tolist())
