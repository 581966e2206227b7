"""zav1_txunit_rd_batch_noalloc (the fixed-scratch variant of the full-RD
transform unit, for callers that cannot allocate) against
zav1_txunit_rd_batch with a caller workspace: identical outputs on two calls
in a row, the second with fewer jobs, and a clear error beyond the scratch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(z, rng, n, ts, w, h, bd, plane, PW, PH):
    mx = (1 << bd) - 1
    jobs = np.zeros(n, z.TXUNIT_RD_JOB)
    preds = []
    for j in range(n):
        x, y = int(rng.integers(0, PW - w + 1)), int(rng.integers(0, PH - h + 1))
        blk = plane[y:y + h, x:x + w].astype(np.int64)
        p = blk + rng.integers(-20 << (bd - 8), (20 << (bd - 8)) + 1, blk.shape)
        preds.append(np.clip(p, 0, mx).astype(np.uint16))
        t = [0, 1, 2, 3, 9, 10, 11][j % 7] if max(w, h) < 32 else 0
        jobs[j] = (x, y, j * w * h, t, w, h, j % 2, j % 13, j % 3, 0, 0, 300 * (j % 5), 61000 + 7 * j)
    return jobs, np.concatenate([p.ravel() for p in preds])


@pytest.mark.parametrize('bd,ts,flags,cut', [(10, 2, 8 | 2, (60, 100)), (8, 3, 1 | 4, (0, 0)), (10, 0, 8, (0, 0))])
def test_noalloc_matches_workspace_variant(gpu, bd, ts, flags, cut):
    z = gpu
    L = z.lib()
    A = z.DeviceArray
    rng = np.random.default_rng(ts * 7 + bd)
    w, h = z.TX_DIMS[ts]
    PH, PW = 256, 320
    mx = (1 << bd) - 1
    yy, xx = np.mgrid[0:PH, 0:PW]
    plane = np.clip((np.sin(xx / 7.0) * np.cos(yy / 5.0) * 0.4 + 0.5) * mx + rng.normal(0, 4 << (bd - 8), (PH, PW)),
                    0, mx).astype(np.uint8 if bd == 8 else np.uint16)
    d_src = A.from_host(plane)
    qidx = 150
    d_costs = A.from_host(z.build_coeff_costs(qidx))
    for n in (1200, 257):                      # the second call smaller than the first
        jobs, pred = _case(z, rng, n, ts, w, h, bd, plane, PW, PH)
        d_pred, d_jobs = A.from_host(pred), A.from_host(jobs.view(np.uint8))
        outs = []
        for noalloc in (False, True):
            d_dist, d_eob, d_rate, d_cul = A(n, np.uint64), A(n, np.int32), A(n, np.int32), A(n, np.uint8)
            d_q = A(n * min(w, 32) * min(h, 32), np.int32)
            d_rec = A(n * w * h, np.uint16)
            if noalloc:
                rc = L.zav1_txunit_rd_batch_noalloc(d_src.ptr, PW, PW, PH, bd, ts, d_pred.ptr, d_pred.nbytes // 2,
                                                    d_jobs.ptr, n, qidx, flags, cut[0], cut[1], d_costs.ptr,
                                                    d_dist.ptr, d_eob.ptr, d_rate.ptr, d_cul.ptr, d_q.ptr,
                                                    d_rec.ptr, None)
                assert rc == 0, z.last_error()
            else:
                d_work = A(z.txunit_rd_workspace(ts, n), np.uint8)
                z.txunit_rd_batch(d_src, PW, PW, PH, bd, ts, d_pred, d_jobs, n, qidx, flags, d_costs, d_work,
                                  d_dist, d_eob, d_rate, d_cul, d_q, d_rec, cut_off=cut)
            z.sync()
            o = [d_dist.to_host(), d_eob.to_host(), d_rate.to_host(), d_cul.to_host(), d_q.to_host()]
            if flags & 6:
                o.append(d_rec.to_host())
            outs.append(o)
        for a, b in zip(*outs):
            assert np.array_equal(a, b), (bd, ts, n)
        assert outs[0][1].min() >= 0


def test_noalloc_rejects_beyond_scratch(gpu):
    z = gpu
    L = z.lib()
    cap = L.zav1_noalloc_scratch_bytes()
    assert cap == 32 << 20
    n = 1
    while z.txunit_rd_workspace(4, n) <= cap:
        n *= 2
    f = ctypes.c_void_p(256)
    rc = L.zav1_txunit_rd_batch_noalloc(f, 64, 64, 64, 8, 4, f, 1 << 30, f, n, 100, 1, 0, 0, f, f, f, f, f, None,
                                        None, None)
    assert rc == z.ZAV1_EINVAL and 'scratch' in z.last_error()
