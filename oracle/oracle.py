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


def _bind_funnel(L):
    L.or_rdcost.argtypes = [ctypes.c_uint64] * 3
    L.or_rdcost.restype = ctypes.c_uint64
    L.or_exchange_sort.argtypes = [_vp, ctypes.c_int, _vp]
    L.or_md_fast_select.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]


def rdcost(lam, rate, dist):
    L = lib()
    _bind_funnel(L)
    return L.or_rdcost(lam, rate, dist)


def exchange_sort(costs):
    """Candidate order after the reference's exchange sort of `costs`."""
    L = lib()
    _bind_funnel(L)
    c = np.ascontiguousarray(costs, dtype=np.uint64)
    idx = np.zeros(len(c), np.int32)
    L.or_exchange_sort(c.ctypes.data, len(c), idx.ctypes.data)
    return idx


def md_fast_select(satd, rate, lam, n_cands, k):
    """(sel [n_units, k] u8, sel_cost [n_units, k] u64) of the MDS0 selection."""
    L = lib()
    _bind_funnel(L)
    s = np.ascontiguousarray(satd, dtype=np.uint32).ravel()
    r = None if rate is None else np.ascontiguousarray(rate, dtype=np.uint32).ravel()
    n_units = s.size // n_cands
    sel = np.zeros((n_units, k), np.uint8)
    cost = np.zeros((n_units, k), np.uint64)
    L.or_md_fast_select(s.ctypes.data, None if r is None else r.ctypes.data, lam, n_units, n_cands, k,
                        sel.ctypes.data, cost.ctypes.data)
    return sel, cost


def md_funnel(src, recon, bd, intra_jobs, n_cands, rate, lam, k, tx_size, tx_type, qindex, use_b=True):
    """MDS0 -> MDS1 chain over units (jobs unit-major).  Returns sel, sel_cost,
    dist, eob, cul, each [n_units, k]."""
    L = lib()
    L.or_md_funnel.argtypes = ([_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                ctypes.c_uint64] + [ctypes.c_int] * 5 + [_vp] * 5)
    s, r = np.ascontiguousarray(src), np.ascontiguousarray(recon)
    j = np.ascontiguousarray(intra_jobs).view(INTRA_JOB)
    nu = len(j) // n_cands
    rt = None if rate is None else np.ascontiguousarray(np.broadcast_to(rate, (nu, n_cands)), dtype=np.uint32)
    sel = np.zeros((nu, k), np.uint8)
    cost = np.zeros((nu, k), np.uint64)
    dist = np.zeros((nu, k), np.uint64)
    eob = np.zeros((nu, k), np.int32)
    cul = np.zeros((nu, k), np.uint8)
    L.or_md_funnel(s.ctypes.data, r.ctypes.data, s.shape[1], bd, j.ctypes.data, nu, n_cands,
                   None if rt is None else rt.ctypes.data, lam, k, tx_size, tx_type, qindex, int(use_b),
                   sel.ctypes.data, cost.ctypes.data, dist.ctypes.data, eob.ctypes.data, cul.ctypes.data)
    return sel, cost, dist, eob, cul


def dr_predict_edges(above, left, w, h, angle, up_a, up_l, bd):
    """Directional prediction from final 160-sample edges (origin 16)."""
    L = lib()
    L.or_dr_predict_edges.argtypes = [_vp, _vp] + [ctypes.c_int] * 6 + [_vp]
    a = np.ascontiguousarray(above, dtype=np.int32)
    lf = np.ascontiguousarray(left, dtype=np.int32)
    out = np.zeros((h, w), np.int32)
    L.or_dr_predict_edges(a.ctypes.data, lf.ctypes.data, w, h, angle, int(up_a), int(up_l), bd, out.ctypes.data)
    return out


def lf_thresholds(level, sharpness):
    L = lib()
    L.or_lf_thresholds.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
    m, li, hv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.or_lf_thresholds(level, sharpness, ctypes.byref(m), ctypes.byref(li), ctypes.byref(hv))
    return m.value, li.value, hv.value


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


class DeblockGeom(ctypes.Structure):
    _fields_ = [('mi_cols', ctypes.c_int), ('mi_rows', ctypes.c_int), ('true_w', ctypes.c_int),
                ('true_h', ctypes.c_int), ('block_id', _vp), ('bw', _vp), ('bh', _vp), ('tw', _vp), ('th', _vp),
                ('skip', _vp)]


def _bind_txunit(L):
    L.or_tx_unit_full.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp] + [ctypes.c_int] * 9 + [_vp, _vp, _vp, _vp,
                                                                                               _vp]


def tx_unit_full(src, pred, tx_size, tx_type, qindex, bd, use_b=False, need_recon=True, spatial=False, crop=None):
    """One transform unit of tx_unit_inner / tx_unit_hbd on a block `src`
    (h x w view of a plane, u8 or u16) with a u16 prediction.  Returns
    (dist, eob, cul, qcoeff packed, recon or None)."""
    L = lib()
    _bind_txunit(L)
    w, h = tx_dims(tx_size)
    s = np.ascontiguousarray(src)
    p = np.ascontiguousarray(pred, dtype=np.uint16)
    pw, ph = min(w, 32), min(h, 32)
    q = np.zeros(pw * ph, np.int32)
    rec = np.zeros((h, w), np.uint16)
    d, e, c = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_uint8()
    cw, ch = crop if crop is not None else (w, h)
    ok = L.or_tx_unit_full(s.ctypes.data, s.shape[1], int(s.dtype == np.uint16), p.ctypes.data, tx_size, tx_type,
                           qindex, bd, int(use_b), int(need_recon), int(spatial), cw, ch, q.ctypes.data,
                           rec.ctypes.data, ctypes.byref(d), ctypes.byref(e), ctypes.byref(c))
    if not ok:
        return None
    return d.value, e.value, c.value, q.reshape(ph, pw), (rec if (need_recon or spatial) else None)


def _bind_lf(L):
    L.or_cdef_constrain.argtypes = [ctypes.c_int] * 3
    L.or_cdef_adjust_strength.argtypes = [ctypes.c_int] * 2
    L.or_cdef_find_dir.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.or_cdef_filter_block.argtypes = [_vp, ctypes.c_int, _vp] + [ctypes.c_int] * 9
    L.or_lpf.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_long] + [ctypes.c_int] * 5
    L.or_deblock_plane.argtypes = [_vp] + [ctypes.c_int] * 9 + [ctypes.POINTER(DeblockGeom)]
    L.or_cdef_search_fb.argtypes = [_vp] * 6 + [ctypes.c_int] * 7 + [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                                                       ctypes.c_int]
    L.or_lf_thresholds.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
    L.or_deblock_search_level.argtypes = [_vp, _vp] + [ctypes.c_int] * 6 + [ctypes.POINTER(DeblockGeom), ctypes.c_int,
                                                                              _vp, _vp, _vp]


def lflib():
    L = lib()
    _bind_lf(L)
    return L


def cdef_find_dir(block8x8, coeff_shift=0):
    b = np.ascontiguousarray(block8x8, dtype=np.uint16)
    v = ctypes.c_int()
    d = lflib().or_cdef_find_dir(b.ctypes.data, b.shape[1], coeff_shift, ctypes.byref(v))
    return d, v.value


def deblock_plane(plane, bd, plane_idx, ss, level_v, level_h, sharpness, geom_arrays, mi_cols, mi_rows,
                  true_w, true_h):
    """In-place deblock of a copy of `plane`; geom_arrays: dict of numpy arrays."""
    p = np.ascontiguousarray(plane).copy()
    g = DeblockGeom(mi_cols, mi_rows, true_w, true_h, geom_arrays['block_id'].ctypes.data,
                    geom_arrays['bw'].ctypes.data, geom_arrays['bh'].ctypes.data, geom_arrays['tw'].ctypes.data,
                    geom_arrays['th'].ctypes.data, geom_arrays['skip'].ctypes.data if 'skip' in geom_arrays else None)
    lflib().or_deblock_plane(p.ctypes.data, p.shape[1], p.shape[1], p.shape[0], bd, plane_idx, ss, level_v,
                             level_h, sharpness, ctypes.byref(g))
    return p


def deblock_search_level(recon, src, bd, plane_idx, ss, sharpness, geom_arrays, mi_cols, mi_rows, true_w, true_h,
                         conv_th=2):
    """search_filter_level with real trials: (best level, sse at level 0, best sse)."""
    r, s_ = np.ascontiguousarray(recon), np.ascontiguousarray(src)
    g = DeblockGeom(mi_cols, mi_rows, true_w, true_h, geom_arrays['block_id'].ctypes.data,
                    geom_arrays['bw'].ctypes.data, geom_arrays['bh'].ctypes.data, geom_arrays['tw'].ctypes.data,
                    geom_arrays['th'].ctypes.data, geom_arrays['skip'].ctypes.data if 'skip' in geom_arrays else None)
    b, e0, be = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    lflib().or_deblock_search_level(r.ctypes.data, s_.ctypes.data, r.shape[1], r.shape[0], bd, plane_idx, ss,
                                    sharpness, ctypes.byref(g), conv_th, ctypes.byref(b), ctypes.byref(e0),
                                    ctypes.byref(be))
    return b.value, e0.value, be.value


def cdef_search_fb(rec, src, width, height, bd, qindex, sub, fbr, fbc, fs, skip=None):
    """rec/src: (y, u, v) tuples of planes (u/v may be None); skip: optional
    per-mi signaled-skip map (mi_rows x mi_cols u8)."""
    sk = np.ascontiguousarray(skip, dtype=np.uint8) if skip is not None else None
    fsa = np.ascontiguousarray(fs, dtype=np.int32)
    my = np.zeros(fsa.size, np.uint64)
    muv = np.zeros(fsa.size, np.uint64)
    ptr = lambda a: a.ctypes.data if a is not None else None
    lflib().or_cdef_search_fb(ptr(rec[0]), ptr(rec[1]), ptr(rec[2]), ptr(src[0]), ptr(src[1]), ptr(src[2]),
                              width, height, bd, qindex, sub, fbr, fbc, fsa.ctypes.data, fsa.size, my.ctypes.data,
                              muv.ctypes.data if rec[1] is not None else None, ptr(sk),
                              sk.shape[1] if sk is not None else 0, sk.shape[0] if sk is not None else 0)
    return my, muv


def _bind_lr(L):
    L.or_count_units.argtypes = [ctypes.c_int, ctypes.c_int]
    L.or_unit_limits.argtypes = [ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_int)] * 4
    L.or_wiener_stats.argtypes = [ctypes.c_int, _vp, _vp] + [ctypes.c_int] * 7 + [_vp, _vp]
    L.or_sse_region.argtypes = [_vp, _vp] + [ctypes.c_int] * 6
    L.or_sse_region.restype = ctypes.c_uint64
    L.or_wiener_filter_plane.argtypes = [_vp, _vp] + [ctypes.c_int] * 5 + [_vp, _vp]
    L.or_wiener_refine_unit.argtypes = [_vp, _vp] + [ctypes.c_int] * 10 + [_vp, _vp]
    L.or_wiener_refine_unit.restype = ctypes.c_int64
    L.or_sgr_param.argtypes = [ctypes.c_int, ctypes.c_int]
    L.or_sgr_unit_stats.argtypes = [_vp, _vp] + [ctypes.c_int] * 8 + [_vp]


def lrlib():
    L = lib()
    _bind_lr(L)
    return L


def unit_limits(unit_size, pw, ph, ss_y, ux, uy):
    v = [ctypes.c_int() for _ in range(4)]
    lrlib().or_unit_limits(unit_size, pw, ph, ss_y, ux, uy, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def wiener_filter_plane(dgd, bd, unit_size, ss_y, types, filters):
    """or_wiener_filter_plane: types (n_units,) 0/1, filters (n_units, 2, 8)
    int16 (hfilter, vfilter); returns the filtered plane."""
    d = np.ascontiguousarray(dgd)
    out = np.zeros_like(d)
    t = np.ascontiguousarray(types, np.uint8)
    f = np.ascontiguousarray(filters, np.int16)
    lrlib().or_wiener_filter_plane(d.ctypes.data, out.ctypes.data, d.shape[1], d.shape[0], bd, unit_size, ss_y,
                                   t.ctypes.data, f.ctypes.data)
    return out


def wiener_refine_unit(dgd, src, bd, win, lim, hf, vf, use_refinement=True, max_one_step=False):
    """or_wiener_refine_unit: returns (sse, hfilter, vfilter)."""
    d, s_ = np.ascontiguousarray(dgd), np.ascontiguousarray(src)
    h, v = np.array(hf, np.int16), np.array(vf, np.int16)
    e = lrlib().or_wiener_refine_unit(d.ctypes.data, s_.ctypes.data, d.shape[1], d.shape[0], bd, win, *lim,
                                      int(use_refinement), int(max_one_step), h.ctypes.data, v.ctypes.data)
    return int(e), h, v


def wiener_stats(win, dgd, src, bd, lim):
    d = np.ascontiguousarray(dgd)
    s = np.ascontiguousarray(src)
    M = np.zeros(win * win, np.int64)
    H = np.zeros(win ** 4, np.int64)
    lrlib().or_wiener_stats(win, d.ctypes.data, s.ctypes.data, d.shape[1], d.shape[0], bd, *lim, M.ctypes.data,
                            H.ctypes.data)
    return M, H.reshape(win * win, win * win)


def sgr_unit_stats(dgd, src, bd, lim, pset):
    d = np.ascontiguousarray(dgd)
    s = np.ascontiguousarray(src)
    out = np.zeros(5, np.int64)
    lrlib().or_sgr_unit_stats(d.ctypes.data, s.ctypes.data, d.shape[1], d.shape[0], bd, *lim, pset, out.ctypes.data)
    return out


def _bind_rdoq(L):
    L.or_optimize_b.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_optimize_b.restype = ctypes.c_int
    L.or_rdoq_eob_cost.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int]
    L.or_cost_coeffs_txb.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp] + [ctypes.c_int] * 6
    for f in ('or_adjusted_tx_size', 'or_txs_ctx', 'or_eob_multi_size'):
        getattr(L, f).argtypes = [ctypes.c_int]


def optimize_b(tcoeff, qcoeff, dqcoeff, eob, scan_, dequant, txb_costs, eob_costs, rdmult, tx_size, tx_class,
               txb_skip_ctx=0, dc_sign_ctx=0, sharpness=0, cut_off_num=0, cut_off_denum=0):
    """enc/quant.rs:1312 optimize_b on one packed block; returns (q, dq, eob).
    txb_costs: int32[944] (one TxbCosts), eob_costs: int32[22] (one EobCosts)."""
    L = lib()
    _bind_rdoq(L)
    t = np.ascontiguousarray(tcoeff, dtype=np.int32)
    q = np.array(qcoeff, dtype=np.int32)
    d = np.array(dqcoeff, dtype=np.int32)
    sc = np.ascontiguousarray(scan_, dtype=np.uint16)
    dqv = np.ascontiguousarray(dequant, dtype=np.int32)
    tc = np.ascontiguousarray(txb_costs, dtype=np.int32)
    ec = np.ascontiguousarray(eob_costs, dtype=np.int32)
    e = L.or_optimize_b(t.ctypes.data, q.ctypes.data, d.ctypes.data, int(eob), sc.ctypes.data, dqv.ctypes.data,
                        tc.ctypes.data, ec.ctypes.data, int(rdmult), tx_size, tx_class, txb_skip_ctx, dc_sign_ctx,
                        sharpness, cut_off_num, cut_off_denum)
    return q, d, e


def rdoq_eob_cost(eob, txb_costs, eob_costs, tx_class):
    L = lib()
    _bind_rdoq(L)
    tc = np.ascontiguousarray(txb_costs, dtype=np.int32)
    ec = np.ascontiguousarray(eob_costs, dtype=np.int32)
    return L.or_rdoq_eob_cost(eob, ec.ctypes.data, tc.ctypes.data, tx_class)


def rdoq_size_ctx(tx_size):
    """(adjusted tx size, txs_ctx, eob multi size) of coeff_c.rs:69/:119/:63."""
    L = lib()
    _bind_rdoq(L)
    return L.or_adjusted_tx_size(tx_size), L.or_txs_ctx(tx_size), L.or_eob_multi_size(tx_size)


def cost_coeffs_txb(qcoeff, eob, scan_, txb_costs, eob_costs, tx_size, tx_class, plane_type=0, txb_skip_ctx=0,
                    dc_sign_ctx=0, txt_rate=0):
    """lf/coeff_rate.rs:60 cost_coeffs_txb of one packed block (eob > 0)."""
    L = lib()
    _bind_rdoq(L)
    q = np.ascontiguousarray(qcoeff, dtype=np.int32)
    sc = np.ascontiguousarray(scan_, dtype=np.uint16)
    tc = np.ascontiguousarray(txb_costs, dtype=np.int32)
    ec = np.ascontiguousarray(eob_costs, dtype=np.int32)
    return L.or_cost_coeffs_txb(q.ctypes.data, int(eob), sc.ctypes.data, tc.ctypes.data, ec.ctypes.data, tx_size,
                                tx_class, plane_type, txb_skip_ctx, dc_sign_ctx, txt_rate)


TXB_WORDS, EOB_WORDS = 944, 22


def build_coeff_costs(base_qindex):
    """enc/quant.rs:514 build_coeff_cost_tables(base_qindex) from the default
    CDFs: (txb int32[10, 944] indexed txs_ctx*2 + plane, eob int32[7, 2, 22])."""
    L = lib()
    L.or_build_coeff_costs.argtypes = [ctypes.c_int, _vp, _vp]
    txb = np.zeros((10, TXB_WORDS), np.int32)
    eob = np.zeros((7, 2, EOB_WORDS), np.int32)
    L.or_build_coeff_costs(int(base_qindex), txb.ctypes.data, eob.ctypes.data)
    return txb, eob


def cost_symbol(p15):
    L = lib()
    L.or_cost_symbol.argtypes = [ctypes.c_uint]
    return L.or_cost_symbol(int(p15))


def intra_txt_rate(tx_size, intra_dir, tx_type):
    """leaf_funnel/rate_tables.rs:291 txt_rate (intra) from the default CDFs."""
    L = lib()
    L.or_intra_txt_rate.argtypes = [ctypes.c_int] * 3
    return L.or_intra_txt_rate(tx_size, intra_dir, tx_type)


def intra_tx_type_used(tx_size, tx_type):
    L = lib()
    L.or_intra_tx_type_used.argtypes = [ctypes.c_int] * 2
    return bool(L.or_intra_tx_type_used(tx_size, tx_type))


def cdef_apply_frame(y, u, v, bd, damping, fb_str, skip=None):
    """or_cdef_apply_frame: the CDEF frame pass of planes y (h x w) and u, v
    (h/2 x w/2, or None) with per-fb strength pairs fb_str (n_fb x 2) and
    the per-mi skip map (mi_rows x mi_cols, or None); returns
    (out_y, out_u, out_v, [filtered_px, changed_px])."""
    L = lib()
    L.or_cdef_apply_frame.argtypes = [_vp] * 6 + [ctypes.c_int] * 4 + [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
    y = np.ascontiguousarray(y)
    h, w = y.shape
    oy = np.zeros_like(y)
    ou = ov = None
    if u is not None:
        u, v = np.ascontiguousarray(u, y.dtype), np.ascontiguousarray(v, y.dtype)
        ou, ov = np.zeros_like(u), np.zeros_like(v)
    fs = np.ascontiguousarray(fb_str, np.uint8)
    sk = None if skip is None else np.ascontiguousarray(skip, np.uint8)
    mr, mc = (0, 0) if sk is None else sk.shape
    stats = np.zeros(2, np.uint64)
    ptr = lambda a: None if a is None else a.ctypes.data
    L.or_cdef_apply_frame(ptr(y), ptr(u), ptr(v), ptr(oy), ptr(ou), ptr(ov), w, h, bd, damping, fs.ctypes.data,
                          ptr(sk), mc, mr, stats.ctypes.data)
    return oy, ou, ov, stats


def wf_frame(cfg, images, leaf_map=None, levels=False):
    """or_wf_frame over (n_images, height, width) samples with a zav1_wf_cfg
    record (numpy, same layout); returns (recon, decisions[img][row][col]):
    rows / columns of leaves for uniform leaves, of 8x8 cells (each leaf's
    record at its top-left cell) with a partition map (n_images x units x
    256 log2 leaf sizes, the partition pre-pass's layout).  levels=True adds
    the winners' quantized levels: (n_images, height, width) int32, each
    leaf's min(L, 32)^2 raster at its top-left corner."""
    L = lib()
    L.or_wf_frame_levels.argtypes = [_vp, _vp, _vp, _vp, _vp]
    cfg = np.array(cfg, copy=True)
    c = cfg[0]
    n, H, W, leaf = int(c['n_images']), int(c['height']), int(c['width']), int(c['leaf'])
    src = np.ascontiguousarray(images)
    rec = np.zeros_like(src)
    lm = None
    if leaf_map is not None:
        lm = np.ascontiguousarray(leaf_map, np.uint8)
        c['leaf_map'] = lm.ctypes.data
        dec = np.zeros((n, H // 8, W // 8), WF_DECISION)
    else:
        c['leaf_map'] = 0
        dec = np.zeros((n, H // leaf, W // leaf), WF_DECISION)
    cfg[0] = c
    lev = np.zeros((n, H, W), np.int32) if levels else None
    L.or_wf_frame_levels(np.ascontiguousarray(cfg).ctypes.data, src.ctypes.data, rec.ctypes.data, dec.ctypes.data,
                         lev.ctypes.data if levels else None)
    del lm
    return (rec, dec, lev) if levels else (rec, dec)


WF_DECISION = np.dtype([('cand', 'u1'), ('tx_type', 'u1'), ('mode', 'u1'), ('fi_mode', 'u1'), ('eob', '<i4'),
                        ('rate', '<i4'), ('mode_rate', '<i4'), ('dist', '<u8'), ('cost', '<u8')])


# ---- partition pre-pass (pd0.c) ----------------------------------------------
class Pd0Cfg(ctypes.Structure):
    _fields_ = [('level', ctypes.c_int), ('qp', ctypes.c_int), ('qindex', ctypes.c_int), ('lambda_', ctypes.c_uint64),
                ('ires', ctypes.c_int), ('cap_max_block', ctypes.c_int), ('max_tx', ctypes.c_int),
                ('min_sq', ctypes.c_int), ('coeff_rate_est_lvl', ctypes.c_int), ('tile_w', ctypes.c_int),
                ('tile_h', ctypes.c_int), ('split_bits', ctypes.c_int32 * 4), ('none_bits', ctypes.c_int32),
                ('skip0_bits', ctypes.c_int32)]


PD0_NODE = np.dtype([('cost', '<u8'), ('flags', 'u1'), ('pad', 'u1', 7)])
PD0_VISITED, PD0_TESTED, PD0_SPLIT = 1, 2, 8
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags='C_CONTIGUOUS')
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags='C_CONTIGUOUS')


def _bind_pd0(L):
    if getattr(L, '_pd0', False):
        return
    L.or_b64_variance.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u16p]
    L.or_qp_th_scaling.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    L.or_kf_full_lambda_8bit.argtypes = [ctypes.c_int, ctypes.c_int]
    L.or_kf_full_lambda_8bit.restype = ctypes.c_uint32
    L.or_max_block_size_allintra.argtypes = [ctypes.c_int, ctypes.c_int]
    L.or_input_resolution_factor.argtypes = [ctypes.c_long]
    L.or_pd0_detector_demotes.argtypes = [_u16p, ctypes.c_int]
    L.or_lvl6_cost.argtypes = [_u16p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.or_lvl6_cost.restype = ctypes.c_uint64
    L.or_pd0_default_rates.argtypes = [ctypes.POINTER(ctypes.c_int32 * 4), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32)]
    L.or_pd0_block_cost_sb.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Pd0Cfg),
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]
    L.or_pd0_block_cost_sb.restype = ctypes.c_uint64
    L.or_pd0_sb.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Pd0Cfg), _vp, _u8p,
                            ctypes.POINTER(ctypes.c_int)]
    L.or_pd0_frame.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Pd0Cfg), _vp, _u8p,
                               _u8p]
    L._pd0 = True


def pd0lib():
    L = lib()
    _bind_pd0(L)
    return L


def b64_variance(plane, ox, oy):
    p = np.ascontiguousarray(plane, np.uint8)
    v = np.zeros(85, np.uint16)
    pd0lib().or_b64_variance(p, p.shape[1], ox, oy, v)
    return v


def qp_th_scaling(qp):
    a, b = ctypes.c_uint32(), ctypes.c_uint32()
    pd0lib().or_qp_th_scaling(qp, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def pd0_default_rates():
    s, n, k = (ctypes.c_int32 * 4)(), ctypes.c_int32(), ctypes.c_int32()
    pd0lib().or_pd0_default_rates(ctypes.byref(s), ctypes.byref(n), ctypes.byref(k))
    return list(s), n.value, k.value


def pd0_cfg(level, qp, qindex, *, ires=0, cap_max_block=0, max_tx=64, min_sq=8, coeff_rate_est_lvl=1, tile_w=0,
            tile_h=0, lam=None):
    L = pd0lib()
    c = Pd0Cfg()
    c.level, c.qp, c.qindex = level, qp, qindex
    c.lambda_ = int(L.or_kf_full_lambda_8bit(qindex, qp)) if lam is None else lam
    c.ires, c.cap_max_block, c.max_tx, c.min_sq = ires, cap_max_block, max_tx, min_sq
    c.coeff_rate_est_lvl, c.tile_w, c.tile_h = coeff_rate_est_lvl, tile_w, tile_h
    sb, nb, k0 = pd0_default_rates()
    for i in range(4):
        c.split_bits[i] = sb[i]
    c.none_bits, c.skip0_bits = nb, k0
    return c


def pd0_block_cost(plane, cfg, mode, sq, ox, oy, sb_x=0, sb_y=0, subres=255):
    p = np.ascontiguousarray(plane, np.uint8)
    st = ctypes.c_int(subres)
    v = pd0lib().or_pd0_block_cost_sb(p, p.shape[1], sb_x, sb_y, ctypes.byref(cfg), mode, sq, ox, oy,
                                      ctypes.byref(st))
    return int(v), st.value


def pd0_sb(plane, cfg, sb_x=0, sb_y=0):
    p = np.ascontiguousarray(plane, np.uint8)
    nodes = np.zeros(341, PD0_NODE)
    lm = np.zeros(256, np.uint8)
    m = ctypes.c_int()
    rc = pd0lib().or_pd0_sb(p, p.shape[1], sb_x, sb_y, ctypes.byref(cfg), nodes.ctypes.data, lm, ctypes.byref(m))
    assert rc == 0
    return nodes, lm.reshape(16, 16), m.value


def pd0_frame(plane, cfg):
    p = np.ascontiguousarray(plane, np.uint8)
    h, w = p.shape
    n = (w // 64) * (h // 64)
    nodes = np.zeros(n * 341, PD0_NODE)
    lm = np.zeros(n * 256, np.uint8)
    modes = np.zeros(n, np.uint8)
    rc = pd0lib().or_pd0_frame(p, w, w, h, ctypes.byref(cfg), nodes.ctypes.data, lm, modes)
    assert rc == 0
    return nodes.reshape(n, 341), lm.reshape(n, 16, 16), modes


def leaf_sizes(leaf_map):
    """Leaf sizes of one 64x64 unit in coding (Z) order from its 16x16 map of
    log2 sizes per 4x4 cell."""
    out = []

    def walk(sq, x, y):
        l = int(leaf_map[y // 4, x // 4])
        if (1 << l) == sq:
            out.append(sq)
            return
        h = sq // 2
        for i in range(4):
            walk(h, x + (i & 1) * h, y + (i >> 1) * h)

    walk(64, 0, 0)
    return out


# ---- AV1 still decoder (av1dec.c) --------------------------------------------
DEC_BLOCK = np.dtype([('x', '<i4'), ('y', '<i4'), ('size', 'u1'), ('mode', 'u1'), ('fi_mode', 'u1'), ('skip', 'u1'),
                      ('delta', 'i1'), ('tx_type', 'u1'), ('pad', 'u1', 2), ('eob', '<i4'), ('pad2', '<i4'),
                      ('level_off', '<u8')])


class DecInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ('width', 'height', 'bd', 'sb', 'qindex', 'tile_cols', 'tile_rows',
                                              'n_blocks')] + [('n_levels', ctypes.c_int64)]


def av1_decode_still(stream, max_w, max_h, bd):
    """Decode a still of the library's subset: (frame, info, blocks, levels)."""
    L = lib()
    L.or_av1_decode_still.argtypes = [_vp, ctypes.c_size_t, _vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(DecInfo),
                                      _vp, ctypes.c_int, _vp, ctypes.c_size_t]
    data = np.frombuffer(bytes(stream), np.uint8)
    frame = np.zeros(max_w * max_h, np.uint16 if bd > 8 else np.uint8)  # decoded rows are `width` apart
    nb = (max_w // 8) * (max_h // 8)
    blocks = np.zeros(nb, DEC_BLOCK)
    levels = np.zeros(max_w * max_h, np.int32)
    info = DecInfo()
    rc = L.or_av1_decode_still(data.ctypes.data, data.size, frame.ctypes.data, max_w, max_h, ctypes.byref(info),
                               blocks.ctypes.data, nb, levels.ctypes.data, levels.size)
    if rc:
        raise ValueError('or_av1_decode_still: %d' % rc)
    img = frame[:info.width * info.height].reshape(info.height, info.width)
    return img, info, blocks[:info.n_blocks], levels[:info.n_levels]
