"""GPU parity: batched and plain-form transforms, quantizers, SSE and SATD
against the oracle on seeded inputs (bit-exact)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TYPE_COL = [0, 1, 0, 1, 2, 0, 2, 1, 2, 3, 0, 3, 1, 3, 2, 3]
TYPE_ROW = [0, 0, 1, 1, 0, 2, 2, 2, 1, 3, 3, 0, 3, 1, 3, 2]


def valid_types(w, h):
    return [t for t in range(16)
            if not ((TYPE_COL[t] in (1, 2) and h > 16) or (TYPE_ROW[t] in (1, 2) and w > 16))]


@pytest.mark.parametrize('tx_size', list(range(19)))
@pytest.mark.parametrize('amp', [255, 1023])
def test_fwd_txfm2d_batch_bitexact(gpu, oracle, tx_size, amp):
    zav1 = gpu
    w, h = zav1.TX_DIMS[tx_size]
    types = valid_types(w, h)
    rng = np.random.default_rng(tx_size * 7 + amp)
    n_jobs = 64 * len(types) + 3  # ragged last workgroup
    x = rng.integers(-amp, amp + 1, (n_jobs, h, w)).astype(np.int32)
    x[0] = 0
    x[1] = amp
    x[2] = -amp
    ty = np.array([types[j % len(types)] for j in range(n_jobs)], dtype=np.uint8)
    d_in = zav1.DeviceArray.from_host(x)
    d_ty = zav1.DeviceArray.from_host(ty)
    d_out = zav1.DeviceArray((n_jobs, h, w), np.int32)
    zav1.fwd_txfm2d_batch(d_in, d_out, d_ty, tx_size, n_jobs)
    got = d_out.to_host()
    for j in range(n_jobs):
        ref = oracle.fwd_txfm2d(x[j], tx_size, int(ty[j]))
        if not np.array_equal(got[j], ref):
            bad = np.argwhere(got[j] != ref)[:4]
            pytest.fail('tx %d type %d job %d mismatch at %s' % (tx_size, ty[j], j, bad.tolist()))


@pytest.mark.parametrize('tx_size', list(range(19)))
@pytest.mark.parametrize('bd', [8, 10])
def test_inv_txfm2d_batch_bitexact(gpu, oracle, tx_size, bd):
    zav1 = gpu
    w, h = zav1.TX_DIMS[tx_size]
    types = [0] if max(w, h) == 64 else valid_types(w, h)
    rng = np.random.default_rng(tx_size * 11 + bd)
    n_jobs = 32 * len(types) + 5
    amp = 4000 if bd == 8 else 16000
    c = rng.integers(-amp, amp + 1, (n_jobs, h, w)).astype(np.int32)
    c[0] = 0
    c[1, 0, 0] = 1 << 20    # exercises the input clamp
    ty = np.array([types[j % len(types)] for j in range(n_jobs)], dtype=np.uint8)
    d_in = zav1.DeviceArray.from_host(c)
    d_ty = zav1.DeviceArray.from_host(ty)
    d_out = zav1.DeviceArray((n_jobs, h, w), np.int32)
    zav1.inv_txfm2d_batch(d_in, d_out, d_ty, tx_size, n_jobs, bd)
    got = d_out.to_host()
    for j in range(n_jobs):
        ref = oracle.inv_txfm2d(c[j], tx_size, int(ty[j]), bd)
        assert np.array_equal(got[j], ref), (tx_size, int(ty[j]), j)


@pytest.mark.parametrize('bd', [8, 10])
@pytest.mark.parametrize('kind', [0, 1])
@pytest.mark.parametrize('log_scale', [0, 1, 2])
def test_quantize_batch_bitexact(gpu, oracle, bd, kind, log_scale):
    zav1 = gpu
    rng = np.random.default_rng(bd * 10 + kind * 3 + log_scale)
    n, jobs = 256, 97
    amp = 40000 if bd == 10 else 30000
    c = rng.integers(-amp, amp + 1, (jobs, n)).astype(np.int32)
    c[:, :8] = rng.integers(-50, 51, (jobs, 8))  # dead-zone region
    qidx = rng.integers(0, 256, jobs).astype(np.uint8)
    qidx[0], qidx[1] = 0, 255
    d_c = zav1.DeviceArray.from_host(c)
    d_q = zav1.DeviceArray((jobs, n), np.int32)
    d_dq = zav1.DeviceArray((jobs, n), np.int32)
    d_qi = zav1.DeviceArray.from_host(qidx)
    zav1.quantize_batch(d_c, d_q, d_dq, d_qi, n, jobs, log_scale, bd, kind)
    q, dq = d_q.to_host(), d_dq.to_host()
    for j in range(jobs):
        row = oracle.quant_row(int(qidx[j]), bd)
        rq, rdq = oracle.quantize('fp' if kind == 0 else 'b', c[j], row, log_scale, bd)
        assert np.array_equal(q[j], rq) and np.array_equal(dq[j], rdq), (j, int(qidx[j]))


def test_sse_batch_bitexact(gpu, oracle):
    zav1 = gpu
    rng = np.random.default_rng(5)
    for n in (16, 64, 1024, 4096):
        jobs = 33
        a = rng.integers(-2 ** 20, 2 ** 20, (jobs, n)).astype(np.int32)
        b = rng.integers(-2 ** 20, 2 ** 20, (jobs, n)).astype(np.int32)
        a[0, 0], b[0, 0] = 2 ** 31 - 1, -(2 ** 31)  # wrapping square
        da, db = zav1.DeviceArray.from_host(a), zav1.DeviceArray.from_host(b)
        out = zav1.DeviceArray(jobs, np.uint64)
        zav1.sse_i32_batch(da, db, out, n, jobs)
        got = out.to_host()
        zav1.sse_i32_batch(da, None, out, n, jobs)
        got_sq = out.to_host()
        for j in range(jobs):
            assert int(got[j]) == oracle.sse_i32(a[j], b[j])
            assert int(got_sq[j]) == oracle.sq_sum_i32(a[j])


@pytest.mark.parametrize('n', [4, 8, 16, 32])
@pytest.mark.parametrize('amp', [255, 1023])
def test_satd_batch_bitexact(gpu, oracle, n, amp):
    """bd8 and bd10 residual ranges; bd10 exercises the 16-bit wrap of the
    16x16/32x32 combine stages."""
    zav1 = gpu
    rng = np.random.default_rng(n + amp)
    jobs = 70
    r = rng.integers(-amp, amp + 1, (jobs, n, n)).astype(np.int16)
    r[0] = amp
    r[1] = -amp
    d_r = zav1.DeviceArray.from_host(r)
    d_c = zav1.DeviceArray((jobs, n * n), np.int32)
    d_s = zav1.DeviceArray(jobs, np.int32)
    zav1.satd_batch(d_r, d_c, d_s, n, jobs)
    coeff, satd = d_c.to_host(), d_s.to_host()
    for j in range(jobs):
        ref = oracle.hadamard(n, r[j])
        assert np.array_equal(coeff[j], ref), (n, j)
        assert int(satd[j]) == int(np.sum(np.abs(ref.astype(np.int64)))), (n, j)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def test_plain_table_matches_oracle(gpu, oracle):
    """The per-call function table (host pointers, one launch per call)."""
    zav1 = gpu
    tab = zav1.dsp_table()
    assert tab.abi_version == 3 and tab.backend == 1
    rng = np.random.default_rng(11)
    i32 = ctypes.c_int32
    for tx_size in (0, 2, 5, 9, 13, 16, 4):
        w, h = zav1.TX_DIMS[tx_size]
        stride = w + 3
        src = rng.integers(-255, 256, (h, stride)).astype(np.int32)
        for t in (valid_types(w, h)[:4]):
            out = np.zeros((h, w), np.int32)
            rc = tab.fwd_txfm2d(_ptr(src, i32), _ptr(out, i32), stride, tx_size, t)
            assert rc == 1
            assert np.array_equal(out, oracle.fwd_txfm2d(src[:, :w].copy(), tx_size, t))
        if max(w, h) == 64:
            out = np.zeros((h, w), np.int32)
            assert tab.fwd_txfm2d(_ptr(src, i32), _ptr(out, i32), stride, tx_size, 1) == 0  # unsupported
    # inverse with stride
    w, h = 16, 16
    c = rng.integers(-3000, 3000, (h, w)).astype(np.int32)
    out = np.zeros((h, w), np.int32)
    assert tab.inv_txfm2d_bd(_ptr(c, i32), _ptr(out, i32), w, 2, 3, 10) == 1
    assert np.array_equal(out, oracle.inv_txfm2d(c, 2, 3, 10))
    # quantizers with the reference's pre-shifted arguments
    row = oracle.quant_row(120, 8)
    coeff = rng.integers(-20000, 20000, 256).astype(np.int32)
    for ls in (0, 1):
        rounding = np.array([(row.round_fp[i] + ((1 << ls) >> 1)) >> ls for i in range(2)], np.int32)
        qfp = np.array(list(row.quant_fp), np.int32)
        deq = np.array(list(row.dequant), np.int32)
        q = np.zeros(256, np.int32)
        dq = np.zeros(256, np.int32)
        assert tab.quantize_fp_raster(_ptr(coeff, i32), _ptr(q, i32), _ptr(dq, i32), 256, _ptr(rounding, i32),
                                      _ptr(qfp, i32), _ptr(deq, i32), ls) == 0
        rq, rdq = oracle.quantize('fp', coeff, row, ls, 8)
        assert np.array_equal(q, rq) and np.array_equal(dq, rdq)
        zb = np.array([(row.zbin[i] + ((1 << ls) >> 1)) >> ls for i in range(2)], np.int32)
        rd = np.array([(row.round[i] + ((1 << ls) >> 1)) >> ls for i in range(2)], np.int32)
        qu = np.array(list(row.quant), np.int32)
        qs = np.array(list(row.quant_shift), np.int32)
        assert tab.quantize_b_raster(_ptr(coeff, i32), _ptr(q, i32), _ptr(dq, i32), 256, _ptr(zb, i32),
                                     _ptr(rd, i32), _ptr(qu, i32), _ptr(qs, i32), _ptr(deq, i32), ls) == 0
        rq, rdq = oracle.quantize('b', coeff, row, ls, 8)
        assert np.array_equal(q, rq) and np.array_equal(dq, rdq)
    a = rng.integers(-5000, 5000, 1000).astype(np.int32)
    b = rng.integers(-5000, 5000, 1000).astype(np.int32)
    res = ctypes.c_uint64()
    assert tab.sse_i32(_ptr(a, i32), _ptr(b, i32), 1000, ctypes.byref(res)) == 0
    assert res.value == oracle.sse_i32(a, b)
    r = rng.integers(-255, 256, (32, 40)).astype(np.int16)
    sat = ctypes.c_int32()
    co = np.zeros(1024, np.int32)
    assert tab.hadamard_satd(_ptr(r, ctypes.c_int16), 40, 32, _ptr(co, i32), ctypes.byref(sat)) == 0
    ref = oracle.hadamard(32, r[:, :32].copy())
    assert np.array_equal(co, ref) and sat.value == int(np.abs(ref).sum())
