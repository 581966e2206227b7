"""ctypes binding of the oracle (CPU restatement of zenav1-svt's DSP kernels).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS')
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags='C_CONTIGUOUS')
_vp = ctypes.c_void_p


def build(force=False):
    """Compile liboracle.so from the C sources beside this file."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.check_call(['make', '-s', '-C', HERE, 'liboracle.so'])


class QuantRow(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32 * 2) for n in
                ('zbin', 'round', 'quant', 'quant_shift', 'round_fp', 'quant_fp', 'dequant')]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_fwd_txfm1d.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int]
        L.or_inv_txfm1d.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int]
        L.or_fwd_txfm2d.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_inv_txfm2d.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_tx_size_dims.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.or_cospi_row.restype = ctypes.POINTER(ctypes.c_int32)
        L.or_sinpi_row.restype = ctypes.POINTER(ctypes.c_int32)
        L.or_build_quant_row.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QuantRow)]
        qargs = [_i32p, ctypes.c_int, ctypes.POINTER(QuantRow), ctypes.c_int, ctypes.c_int, _i32p, _i32p]
        L.or_quantize_fp.argtypes = qargs
        L.or_quantize_b.argtypes = qargs
        L.or_sse_i32.argtypes = [_i32p, _i32p, ctypes.c_int]
        L.or_sse_i32.restype = ctypes.c_uint64
        L.or_sq_sum_i32.argtypes = [_i32p, ctypes.c_int]
        L.or_sq_sum_i32.restype = ctypes.c_uint64
        L.or_sse_pix.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_sse_pix.restype = ctypes.c_uint64
        L.or_hadamard.argtypes = [ctypes.c_int, _i16p, ctypes.c_int, _i32p]
        L.or_satd.argtypes = [_i32p, ctypes.c_int]
        L.or_satd.restype = ctypes.c_int32
        L.or_block_satd.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_block_satd.restype = ctypes.c_uint64
        _lib = L
    return _lib


TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
           (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def tx_dims(tx_size):
    return TX_DIMS[tx_size]


def fwd_txfm1d(kind, vec, cos_bit):
    a = np.ascontiguousarray(vec, dtype=np.int32)
    out = np.zeros_like(a)
    ok = lib().or_fwd_txfm1d(kind, a.size, a, out, cos_bit)
    return out if ok else None


def inv_txfm1d(kind, vec, rng):
    a = np.ascontiguousarray(vec, dtype=np.int32)
    out = np.zeros_like(a)
    ok = lib().or_inv_txfm1d(kind, a.size, a, out, rng)
    return out if ok else None


def fwd_txfm2d(block, tx_size, tx_type):
    """block: (h, w) int32 residual.  Returns (h, w) coefficients or None."""
    w, h = tx_dims(tx_size)
    a = np.ascontiguousarray(block, dtype=np.int32).reshape(h, w)
    out = np.zeros((h, w), dtype=np.int32)
    ok = lib().or_fwd_txfm2d(a.ravel(), out.ravel(), w, tx_size, tx_type)
    return out if ok else None


def inv_txfm2d(coeff, tx_size, tx_type, bd):
    w, h = tx_dims(tx_size)
    a = np.ascontiguousarray(coeff, dtype=np.int32).reshape(h, w)
    out = np.zeros((h, w), dtype=np.int32)
    ok = lib().or_inv_txfm2d(a.ravel(), out.ravel(), w, tx_size, tx_type, bd)
    return out if ok else None


def quant_row(qindex, bd):
    r = QuantRow()
    lib().or_build_quant_row(qindex, bd, ctypes.byref(r))
    return r


def quantize(kind, coeff, row, log_scale, bd):
    c = np.ascontiguousarray(coeff, dtype=np.int32).ravel()
    q = np.zeros_like(c)
    dq = np.zeros_like(c)
    fn = lib().or_quantize_fp if kind == 'fp' else lib().or_quantize_b
    fn(c, c.size, ctypes.byref(row), log_scale, bd, q, dq)
    return q, dq


def sse_i32(a, b):
    a = np.ascontiguousarray(a, dtype=np.int32).ravel()
    b = np.ascontiguousarray(b, dtype=np.int32).ravel()
    return lib().or_sse_i32(a, b, a.size)


def sq_sum_i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32).ravel()
    return lib().or_sq_sum_i32(a, a.size)


def hadamard(n, res):
    r = np.ascontiguousarray(res, dtype=np.int16).reshape(n, n)
    out = np.zeros(n * n, dtype=np.int32)
    lib().or_hadamard(n, r.ravel(), n, out)
    return out


def block_satd(src, pred, hbd):
    s = np.ascontiguousarray(src)
    p = np.ascontiguousarray(pred)
    h, w = s.shape
    return lib().or_block_satd(s.ctypes.data, w, p.ctypes.data, w, w, h, int(hbd))


def _bind_txsearch(L):
    L.or_scan.argtypes = [ctypes.c_int, ctypes.c_int, np.ctypeslib.ndpointer(dtype=np.uint16, flags='C_CONTIGUOUS')]
    L.or_scan_class.argtypes = [ctypes.c_int]
    L.or_tx_unit.argtypes = [_i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_uint32)]
    L.or_txsearch_plane.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]


def scan(tx_size, scan_class):
    L = lib()
    _bind_txsearch(L)
    buf = np.zeros(1024, np.uint16)
    n = L.or_scan(tx_size, scan_class, buf)
    return buf[:n].copy()


def tx_unit(residual, tx_size, tx_type, qindex, bd, use_b=False):
    L = lib()
    _bind_txsearch(L)
    r = np.ascontiguousarray(residual, dtype=np.int32).ravel()
    d, e, s = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_uint32()
    ok = L.or_tx_unit(r, tx_size, tx_type, qindex, bd, int(use_b), None, ctypes.byref(d), ctypes.byref(e),
                      ctypes.byref(s))
    return (d.value, e.value, s.value) if ok else None


def txsearch_plane(plane, bd, tx_size, xy, types, qindex):
    """Reference tx-type search over blocks xy (n,2) of a uint16 plane."""
    L = lib()
    _bind_txsearch(L)
    p = np.ascontiguousarray(plane, dtype=np.uint16)
    xy = np.ascontiguousarray(xy, dtype=np.int32)
    t = np.ascontiguousarray(types, dtype=np.uint8)
    nb, nt = xy.shape[0], t.size
    dist = np.zeros(nb * nt, np.uint64)
    eob = np.zeros(nb * nt, np.int32)
    ls = np.zeros(nb * nt, np.uint32)
    L.or_txsearch_plane(p.ctypes.data, p.shape[1], p.shape[1], p.shape[0], bd, tx_size, xy.ctypes.data, nb,
                        t.ctypes.data, nt, qindex, dist.ctypes.data, eob.ctypes.data, ls.ctypes.data)
    return dist, eob, ls


INTRA_JOB = np.dtype([('x', '<i4'), ('y', '<i4'), ('w', 'u1'), ('h', 'u1'), ('mode', 'u1'), ('delta', 'i1'),
                      ('fi_mode', 'u1'), ('edge_filter', 'u1'), ('filt_type', 'u1'), ('pad0', 'u1'),
                      ('n_top_px', '<i2'), ('n_topright_px', '<i2'), ('n_left_px', '<i2'),
                      ('n_bottomleft_px', '<i2'), ('out_off', '<u4'), ('pad1', '<u4')])


def intra_predict(plane, bd, job):
    """Predict one unit; job is a length-1 INTRA_JOB record array."""
    L = lib()
    L.or_intra_predict.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
    p = np.ascontiguousarray(plane)
    j = np.ascontiguousarray(job, dtype=INTRA_JOB).reshape(1)
    out = np.zeros(int(j['w'][0]) * int(j['h'][0]), np.int32)
    ok = L.or_intra_predict(p.ctypes.data, p.shape[1], bd, j.ctypes.data, out.ctypes.data)
    return out.reshape(int(j['h'][0]), int(j['w'][0])) if ok else None


def cfl_predict(luma, bd, cw, ch, dc, alpha_q3):
    L = lib()
    L.or_cfl_predict.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp]
    lm = np.ascontiguousarray(luma)
    d = np.ascontiguousarray(dc, dtype=np.int32)
    out = np.zeros(cw * ch, np.int32)
    L.or_cfl_predict(lm.ctypes.data, lm.shape[1], bd, cw, ch, d.ctypes.data, alpha_q3, out.ctypes.data)
    return out.reshape(ch, cw)
