"""GPU parity of the batched intra predictors and CfL against the oracle
(every mode, angle delta, filter-intra mode, edge filter / upsample and
availability pattern; 8- and 10-bit; bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
         (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def make_jobs(zav1, rng, n):
    jobs = np.zeros(n, zav1.INTRA_JOB)
    off = 0
    for i in range(n):
        w, h = SIZES[rng.integers(0, len(SIZES))]
        mode = int(rng.integers(0, 13))
        fi = 5
        if rng.uniform() < 0.15 and w <= 32 and h <= 32:
            fi = int(rng.integers(0, 5))
            mode = 0
        delta = int(rng.integers(-3, 4)) if 1 <= mode <= 8 else 0
        avail = rng.uniform()
        nt = w if avail < 0.7 else int(rng.integers(0, w + 1))
        nl = h if avail < 0.7 or avail > 0.85 else int(rng.integers(0, h + 1))
        jobs[i] = (int(rng.integers(72, 184)), int(rng.integers(72, 184)), w, h, mode, delta, fi,
                   int(rng.uniform() < 0.6), int(rng.uniform() < 0.3), 0,
                   nt, int(rng.integers(0, w + 1)) if nt == w else 0,
                   nl, int(rng.integers(0, h + 1)) if nl == h else 0, off, 0)
        off += w * h
    return jobs, off


@pytest.mark.parametrize('bd', [8, 10])
def test_intra_pred_batch_bitexact(gpu, oracle, bd):
    zav1 = gpu
    rng = np.random.default_rng(100 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    # smooth-ish content with edges so interpolation and filters matter
    y, x = np.mgrid[0:320, 0:320]
    plane = ((np.sin(x / 9.0) + np.cos(y / 13.0)) * 0.25 + 0.5) * ((1 << bd) - 1)
    plane = np.clip(plane + rng.normal(0, 8 << (bd - 8), plane.shape), 0, (1 << bd) - 1).astype(dt)
    jobs, total = make_jobs(zav1, rng, 3000)
    d_plane = zav1.DeviceArray.from_host(plane)
    d_jobs = zav1.DeviceArray.from_host(jobs.view(np.uint8))
    d_out = zav1.DeviceArray(total, np.uint16)
    zav1.intra_pred_batch(d_plane, 320, bd, d_jobs, len(jobs), d_out)
    out = d_out.to_host()
    oj = jobs.view(oracle.INTRA_JOB)
    for i in range(len(jobs)):
        ref = oracle.intra_predict(plane, bd, oj[i:i + 1])
        w, h = int(jobs['w'][i]), int(jobs['h'][i])
        got = out[jobs['out_off'][i]:jobs['out_off'][i] + w * h].reshape(h, w)
        if not np.array_equal(got, ref):
            bad = np.argwhere(got != ref)[:3]
            pytest.fail('job %d %s mismatch at %s: got %s ref %s' % (
                i, jobs[i], bad.tolist(), got[tuple(bad[0])], ref[tuple(bad[0])]))


@pytest.mark.parametrize('bd', [8, 10])
def test_cfl_batch_bitexact(gpu, oracle, bd):
    zav1 = gpu
    rng = np.random.default_rng(7 + bd)
    dt = np.uint8 if bd == 8 else np.uint16
    luma = rng.integers(0, 1 << bd, (256, 256)).astype(dt)
    dims = [(4, 4), (8, 8), (16, 16), (32, 32), (4, 8), (8, 4), (16, 32), (32, 8), (4, 16)]
    jobs = np.zeros(400, zav1.CFL_JOB)
    dcs, off = [], 0
    for i in range(400):
        cw, ch = dims[i % len(dims)]
        dc = rng.integers(0, 1 << bd, cw * ch).astype(np.uint16)
        dcs.append(dc)
        jobs[i] = (int(rng.integers(0, 256 - 2 * cw)) & ~1, int(rng.integers(0, 256 - 2 * ch)) & ~1, cw, ch,
                   int(rng.integers(-128, 129)), off, off, 0)
        off += cw * ch
    d_l = zav1.DeviceArray.from_host(luma)
    d_j = zav1.DeviceArray.from_host(jobs.view(np.uint8))
    d_dc = zav1.DeviceArray.from_host(np.concatenate(dcs))
    d_o = zav1.DeviceArray(off, np.uint16)
    zav1.cfl_batch(d_l, 256, bd, d_j, len(jobs), d_dc, d_o)
    out = d_o.to_host()
    for i in range(len(jobs)):
        lx, ly, cw, ch, a = (int(jobs[k][i]) for k in ('lx', 'ly', 'cw', 'ch', 'alpha_q3'))
        sub = np.ascontiguousarray(luma[ly:ly + 2 * ch, lx:lx + 2 * cw])
        ref = oracle.cfl_predict(sub, bd, cw, ch, dcs[i].astype(np.int32).reshape(ch, cw), a)
        got = out[jobs['out_off'][i]:jobs['out_off'][i] + cw * ch].reshape(ch, cw)
        assert np.array_equal(got, ref), (i, cw, ch, a)
