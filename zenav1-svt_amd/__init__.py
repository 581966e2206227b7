"""zav1: MI355X (gfx950) backend for zenav1-svt's intra-encode DSP kernels.

Python host mirror of the C-ABI in include/zav1_dsp.h (ctypes, no torch).
The HIP library must be built (build.py); if it is missing, every entry
point raises — there is no CPU fallback in this package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, 'libzav1_hip.so')
HEADER = os.path.join(ROOT, 'include', 'zav1_dsp.h')

ZAV1_OK = 0
ZAV1_EINVAL = -22
ZAV1_ENOTSUP = -95
ZAV1_ENOMEM = -12
ZAV1_ECANCELED = -125
ZAV1_EIO = -5

TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
           (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


class Zav1Error(RuntimeError):
    def __init__(self, code, what):
        super().__init__('%s failed with %d: %s' % (what, code, last_error()))
        self.code = code


def build():
    from importlib import util
    spec = util.spec_from_file_location('zav1_build', os.path.join(HERE, 'build.py'))
    mod = util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build()


_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_lib = None


class QuantRow(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32 * 2) for n in
                ('zbin', 'round', 'quant', 'quant_shift', 'round_fp', 'quant_fp', 'dequant')]


_I32P = ctypes.POINTER(ctypes.c_int32)
_I16P = ctypes.POINTER(ctypes.c_int16)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_U16P = ctypes.POINTER(ctypes.c_uint16)
_I64P = ctypes.POINTER(ctypes.c_int64)
_sz = ctypes.c_size_t


class DspV1(ctypes.Structure):
    _fields_ = [
        ('abi_version', ctypes.c_uint32),
        ('backend', ctypes.c_uint32),
        ('fwd_txfm2d', ctypes.CFUNCTYPE(_i32, _I32P, _I32P, ctypes.c_size_t, ctypes.c_uint8, ctypes.c_uint8)),
        ('inv_txfm2d_bd', ctypes.CFUNCTYPE(_i32, _I32P, _I32P, ctypes.c_size_t, ctypes.c_uint8, ctypes.c_uint8,
                                           ctypes.c_uint8)),
        ('quantize_fp_raster', ctypes.CFUNCTYPE(_i32, _I32P, _I32P, _I32P, ctypes.c_size_t, _I32P, _I32P, _I32P,
                                                ctypes.c_int32)),
        ('quantize_b_raster', ctypes.CFUNCTYPE(_i32, _I32P, _I32P, _I32P, ctypes.c_size_t, _I32P, _I32P, _I32P,
                                               _I32P, _I32P, ctypes.c_int32)),
        ('sse_i32', ctypes.CFUNCTYPE(_i32, _I32P, _I32P, ctypes.c_size_t, _U64P)),
        ('hadamard_satd', ctypes.CFUNCTYPE(_i32, _I16P, ctypes.c_size_t, _i32, _I32P, _I32P)),
        # ABI 2
        ('residual_i32', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _sz, _sz, _sz, _I32P)),
        ('residual_i32_u16', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _sz, _sz, _sz, _I32P)),
        ('recon_add_clamp', ctypes.CFUNCTYPE(_i32, _vp, _sz, _I32P, _sz, _sz, _vp)),
        ('recon_add_clamp_u16', ctypes.CFUNCTYPE(_i32, _vp, _sz, _I32P, _sz, _sz, _vp, _i32)),
        ('sse_u8', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _sz, _sz, _sz, _U64P)),
        ('sse_u16', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _sz, _sz, _sz, _U64P)),
        ('intra_pred', ctypes.CFUNCTYPE(_i32, _i32, _vp, _sz, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32)),
        ('dr_predictor_edged', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _vp, _sz, _sz, _i32, _i32, _i32, _i32, _i32,
                                                _i32)),
        ('predict_filter_intra', ctypes.CFUNCTYPE(_i32, _vp, _sz, _vp, _vp, _i32, _i32, _i32, _i32)),
        ('cfl_predict', ctypes.CFUNCTYPE(_i32, _I16P, _vp, _sz, _vp, _sz, _i32, _i32, _i32, _i32)),
        ('lpf', ctypes.CFUNCTYPE(_i32, _vp, _sz, _sz, _sz, _i32, _i32, ctypes.c_uint8, ctypes.c_uint8,
                                 ctypes.c_uint8, _i32)),
        ('cdef_find_dir', ctypes.CFUNCTYPE(_i32, _U16P, _sz, _i32, _I32P)),
        ('cdef_filter_block', ctypes.CFUNCTYPE(_i32, _vp, _sz, _U16P, _sz, _sz, _i32, _i32, _i32, _i32, _i32, _i32,
                                               _i32, _sz, _i32)),
        ('compute_stats', ctypes.CFUNCTYPE(_i32, _i32, _vp, _sz, _sz, _sz, _vp, _sz, _sz, _sz, _i32, _i32, _i32,
                                           _i32, _I64P, _I64P, _i32)),
        # ABI 3
        ('quantize_fp_hbd', ctypes.CFUNCTYPE(_i32, _I32P, _U16P, _sz, ctypes.POINTER(QuantRow), ctypes.c_int32,
                                             _I32P, _I32P)),
        ('quantize_b_hbd', ctypes.CFUNCTYPE(_i32, _I32P, _U16P, _sz, ctypes.POINTER(QuantRow), ctypes.c_int32,
                                            _I32P, _I32P)),
        ('cdef_dist', ctypes.CFUNCTYPE(_i32, _vp, _sz, _sz, _sz, _vp, _sz, _vp, _sz, _i32, _i32, _sz, _i32, _U64P)),
    ]


# (name, argtypes) of every exported C-ABI symbol; mirrors include/zav1_dsp.h.
EXPORTS = {
    'zav1_init': [_i32],
    'zav1_device_count': [],
    'zav1_malloc': [ctypes.POINTER(_vp), ctypes.c_size_t],
    'zav1_free': [_vp],
    'zav1_host_alloc': [ctypes.POINTER(_vp), ctypes.c_size_t],
    'zav1_host_free': [_vp],
    'zav1_memcpy_h2d': [_vp, _vp, ctypes.c_size_t, _vp],
    'zav1_memcpy_d2h': [_vp, _vp, ctypes.c_size_t, _vp],
    'zav1_memcpy_d2d': [_vp, _vp, ctypes.c_size_t, _vp],
    'zav1_memset': [_vp, _i32, ctypes.c_size_t, _vp],
    'zav1_stream_create': [ctypes.POINTER(_vp)],
    'zav1_stream_destroy': [_vp],
    'zav1_stream_sync': [_vp],
    'zav1_device_sync': [],
    'zav1_event_create': [ctypes.POINTER(_vp)],
    'zav1_event_destroy': [_vp],
    'zav1_event_record': [_vp, _vp],
    'zav1_event_elapsed_ms': [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    'zav1_stream_wait_event': [_vp, _vp],
    'zav1_event_query': [_vp],
    'zav1_grid_xy': [_vp, _i32, _i32, _i32, _i32, _vp],
    'zav1_last_error': [],
    'zav1_build_quant_row': [_i32, _i32, ctypes.POINTER(QuantRow)],
    'zav1_build_quant_row_sharp': [_i32, _i32, _i32, ctypes.POINTER(QuantRow)],
    'zav1_dsp_get': [_i32],
    'zav1_fwd_txfm2d_batch': [_vp, _vp, _vp, _i32, _i32, _vp],
    'zav1_inv_txfm2d_batch': [_vp, _vp, _vp, _i32, _i32, _i32, _vp],
    'zav1_quantize_batch': [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp],
    'zav1_coeff_rate_batch': [_i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp],
    'zav1_rdoq_lane_min': [_i32, _i32],
    'zav1_rdoq_workspace': [_i32, _i32],
    'zav1_rdoq_batch_ws': [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, ctypes.c_uint32,
                           ctypes.c_uint32, _vp, ctypes.c_size_t, _vp],
    'zav1_cdef_apply': [_vp] * 6 + [_i32] * 4 + [_vp, _vp, _i32, _i32, _vp, _vp],
    'zav1_wiener_apply': [_vp, _vp] + [_i32] * 5 + [_vp, _vp],
    'zav1_wiener_refine': [_vp, _vp] + [_i32] * 8 + [_vp, _vp, _vp],
    'zav1_wf_run_timed': [_vp, _vp, _vp, _i32],
    'zav1_wf_autotune': [_vp, _i32, _i32, _vp, _vp, _vp],
    'zav1_rdoq_batch': [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                        _vp],
    'zav1_sse_i32_batch': [_vp, _vp, _vp, _i32, _i32, _vp],
    'zav1_satd_batch': [_vp, _vp, _vp, _i32, _i32, _vp],
    'zav1_scan_order': [_i32, _i32, _vp],
    'zav1_txsearch_batch': [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _vp, _vp],
    'zav1_txsearch_batch_q': [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp],
    'zav1_intra_pred_batch': [_vp, _i32, _i32, _vp, _i32, _vp, _vp],
    'zav1_cfl_batch': [_vp, _i32, _i32, _vp, _i32, _vp, _vp, _vp],
    'zav1_txunit_batch': [_vp, _i32, _i32, _i32, _i32, _i32, _vp, ctypes.c_size_t, _vp, _i32, _i32, ctypes.c_uint,
                          _vp, _vp, _vp, _vp, _vp, _vp],
    'zav1_pred_satd_batch': [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, ctypes.c_size_t, _vp, _i32, _vp, _vp],
    'zav1_md_fast_select': [_vp, _vp, ctypes.c_uint64, _i32, _i32, _i32, _vp, _i32, _i32, _i32, _vp, _i32, _i32, _vp,
                            _vp, _vp, _vp],
    'zav1_md_fast_satd': [_vp, _vp] + [_i32] * 6 + [_vp, _i32, _vp, _i32, _vp, _vp],
    'zav1_ctx_create': [_i32, ctypes.c_size_t, ctypes.POINTER(_vp)],
    'zav1_ctx_destroy': [_vp],
    'zav1_ctx_alloc': [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)],
    'zav1_ctx_usage': [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    'zav1_ctx_reset': [_vp],
    'zav1_ctx_stream': [_vp],
    'zav1_image_upload': [_vp, _vp, ctypes.c_size_t, _i32, _i32, _i32, ctypes.POINTER(_vp)],
    'zav1_txsearch_frame': [_vp, _vp, _i32, _i32, _i32, _vp, _i32, _vp],
    'zav1_md_wave_submit': [_vp, _vp],
    'zav1_wave_collect': [_vp],
    'zav1_md_predict_survivors': [_vp] + [_i32] * 6 + [_vp, _i32, _vp, _i32, _vp, _i32, _vp, _vp],
    'zav1_deblock_plane': [_vp] + [_i32] * 9 + [_vp, _vp],
    'zav1_deblock_level_sse': [_vp, _vp] + [_i32] * 6 + [_vp, _vp, _i32, _vp, ctypes.c_size_t, _vp, _vp],
    'zav1_deblock_search_level': [_vp, _vp] + [_i32] * 6 + [_vp, _i32, _vp, ctypes.c_size_t, _vp,
                                                            ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_int64),
                                                            ctypes.POINTER(ctypes.c_int64), _vp],
    'zav1_cdef_search': [_vp] * 6 + [_i32] * 5 + [_vp, _i32, _i32, _vp, _vp, _vp, _vp],
    'zav1_rest_unit_grid': [_i32, _i32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    'zav1_wiener_stats': [_vp, _vp] + [_i32] * 6 + [_vp, _vp, _vp],
    'zav1_unit_sse': [_vp, _vp] + [_i32] * 5 + [_vp, _vp],
    'zav1_sgr_stats': [_vp, _vp] + [_i32] * 5 + [_vp, _vp],
    'zav1_build_coeff_costs': [_i32, _vp],
    'zav1_intra_txt_rate': [_i32, _i32, _i32],
    'zav1_txunit_rd_workspace': [_i32, _i32, ctypes.POINTER(ctypes.c_size_t)],
    'zav1_txunit_rd_batch': [_vp, _i32, _i32, _i32, _i32, _i32, _vp, ctypes.c_size_t, _vp, _i32, _i32, ctypes.c_uint,
                             ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp,
                             _vp, _vp],
    'zav1_txunit_rd_batch_noalloc': [_vp, _i32, _i32, _i32, _i32, _i32, _vp, ctypes.c_size_t, _vp, _i32, _i32,
                                     ctypes.c_uint, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp],
    'zav1_noalloc_scratch_bytes': [],
    'zav1_intra_mode_rate': [_i32] * 7,
    'zav1_kf_full_lambda': [_i32, _i32, _i32],
    'zav1_wf_footprint': [_vp, ctypes.POINTER(ctypes.c_size_t)],
    'zav1_wf_create': [_vp, ctypes.POINTER(_vp)],
    'zav1_wf_info': [_vp, _vp],
    'zav1_wf_run': [_vp, _i32, _vp],
    'zav1_wf_run_steps': [_vp, _i32, _i32, _vp],
    'zav1_wf_run_plain': [_vp, _i32, _i32, _vp],
    'zav1_wf_destroy': [_vp],
    'zav1_wf_last_error': [],
    'zav1_pd0_rates': [_vp],
    'zav1_pd0_nodes_per_unit': [_vp],
    'zav1_pd0_run': [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'zav1_write_still': [_vp, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)],
    'zav1_wf_write_stills': [_vp, _i32, _i32, _i32, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _vp],
}

# numpy layouts of the job structs (match include/zav1_dsp.h)
INTRA_JOB = np.dtype([('x', '<i4'), ('y', '<i4'), ('w', 'u1'), ('h', 'u1'), ('mode', 'u1'), ('delta', 'i1'),
                      ('fi_mode', 'u1'), ('edge_filter', 'u1'), ('filt_type', 'u1'), ('pad0', 'u1'),
                      ('n_top_px', '<i2'), ('n_topright_px', '<i2'), ('n_left_px', '<i2'),
                      ('n_bottomleft_px', '<i2'), ('out_off', '<u4'), ('pad1', '<u4')])
TXUNIT_JOB = np.dtype([('x', '<i4'), ('y', '<i4'), ('pred_off', '<u4'), ('tx_type', 'u1'), ('crop_w', 'u1'),
                       ('crop_h', 'u1'), ('pad', 'u1')])
TXU_QUANT_B, TXU_SPATIAL, TXU_RECON, TXU_RDOQ = 1, 2, 4, 8
TXUNIT_RD_JOB = np.dtype([('x', '<i4'), ('y', '<i4'), ('pred_off', '<u4'), ('tx_type', 'u1'), ('crop_w', 'u1'),
                          ('crop_h', 'u1'), ('plane_type', 'u1'), ('txb_skip_ctx', 'u1'), ('dc_sign_ctx', 'u1'),
                          ('sharpness', 'u1'), ('pad', 'u1'), ('txt_rate', '<i4'), ('rdmult', '<i8')])
CFL_JOB = np.dtype([('lx', '<i4'), ('ly', '<i4'), ('cw', '<i4'), ('ch', '<i4'), ('alpha_q3', '<i4'),
                    ('dc_off', '<u4'), ('out_off', '<u4'), ('pad', '<u4')])


def lib():
    """Load libzav1_hip.so (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('zav1: %s is not built (run zenav1-svt_amd/build.py); '
                               'there is no CPU fallback' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in EXPORTS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i32
        L.zav1_last_error.restype = ctypes.c_char_p
        L.zav1_dsp_get.restype = ctypes.POINTER(DspV1)
        L.zav1_ctx_stream.restype = _vp
        L.zav1_noalloc_scratch_bytes.restype = ctypes.c_size_t
        L.zav1_wf_last_error.restype = ctypes.c_char_p
        L.zav1_kf_full_lambda.restype = ctypes.c_uint64
        L.zav1_rdoq_workspace.restype = ctypes.c_size_t
        _lib = L
    return _lib


def last_error():
    try:
        return lib().zav1_last_error().decode()
    except Exception:  # noqa: BLE001 - diagnostics only
        return ''


def check(rc, what):
    if rc < 0:
        raise Zav1Error(rc, what)
    return rc


def init(device=0):
    check(lib().zav1_init(device), 'zav1_init')


def device_count():
    return lib().zav1_device_count()


class DeviceArray:
    """A device allocation with a numpy dtype/shape, owned by Python."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib().zav1_malloc(ctypes.byref(p), self.nbytes), 'zav1_malloc')
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr, stream)
        return d

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(lib().zav1_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes, stream), 'memcpy_h2d')
        check(lib().zav1_stream_sync(stream), 'stream_sync')

    def to_host(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().zav1_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream), 'memcpy_d2h')
        return out

    def zero(self, stream=None):
        check(lib().zav1_memset(self.ptr, 0, self.nbytes, stream), 'memset')

    def free(self):
        if self.ptr:
            lib().zav1_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter teardown
            pass


def sync(stream=None):
    check(lib().zav1_stream_sync(stream), 'stream_sync')


def stream_create():
    s = ctypes.c_void_p()
    check(lib().zav1_stream_create(ctypes.byref(s)), 'stream_create')
    return s


def stream_sync(s):
    check(lib().zav1_stream_sync(s), 'stream_sync')


def stream_destroy(s):
    check(lib().zav1_stream_destroy(s), 'stream_destroy')


def event_create():
    e = ctypes.c_void_p()
    check(lib().zav1_event_create(ctypes.byref(e)), 'event_create')
    return e.value


def event_record(e, stream):
    check(lib().zav1_event_record(e, stream), 'event_record')


def event_query(e):
    """True once the event's recorded work has completed."""
    return check(lib().zav1_event_query(e), 'event_query') == 1


def event_destroy(e):
    check(lib().zav1_event_destroy(e), 'event_destroy')


def quant_row(qindex, bd):
    r = QuantRow()
    check(lib().zav1_build_quant_row(qindex, bd, ctypes.byref(r)), 'build_quant_row')
    return r


def dsp_table():
    p = lib().zav1_dsp_get(1)
    if not p:
        raise RuntimeError('zav1: no usable GPU for the HIP backend')
    return p.contents


# ---- batched wrappers (device arrays) --------------------------------------
def fwd_txfm2d_batch(d_in, d_out, d_types, tx_size, n_jobs, stream=None):
    check(lib().zav1_fwd_txfm2d_batch(d_in.ptr, d_out.ptr, d_types.ptr, tx_size, n_jobs, stream),
          'fwd_txfm2d_batch')


def inv_txfm2d_batch(d_in, d_out, d_types, tx_size, n_jobs, bd, stream=None):
    check(lib().zav1_inv_txfm2d_batch(d_in.ptr, d_out.ptr, d_types.ptr, tx_size, n_jobs, bd, stream),
          'inv_txfm2d_batch')


def quantize_batch(d_coeff, d_q, d_dq, d_qidx, n_per_job, n_jobs, log_scale, bd, kind, stream=None):
    check(lib().zav1_quantize_batch(d_coeff.ptr, d_q.ptr, d_dq.ptr, d_qidx.ptr, n_per_job, n_jobs,
                                    log_scale, bd, kind, stream), 'quantize_batch')


# zav1_rdoq_job / zav1_coeff_costs as numpy layouts (include/zav1_dsp.h)
RDOQ_JOB = np.dtype([('rdmult', '<i8'), ('coeff', '<u4'), ('tx_class', 'u1'), ('plane_type', 'u1'),
                     ('txb_skip_ctx', 'u1'), ('dc_sign_ctx', 'u1'), ('sharpness', 'u1'), ('pad', 'u1', 7)])
TXB_COST_WORDS, EOB_COST_WORDS = 944, 22
COEFF_COSTS_WORDS = 10 * TXB_COST_WORDS + 14 * EOB_COST_WORDS


RATE_JOB = np.dtype([('coeff', '<u4'), ('txt_rate', '<i4'), ('tx_class', 'u1'), ('plane_type', 'u1'),
                     ('txb_skip_ctx', 'u1'), ('dc_sign_ctx', 'u1')])


def coeff_rate_batch(tx_size, d_qcoeff, d_eob, d_jobs, n_jobs, d_costs, d_scans, d_rate, stream=None):
    """Coefficient rate of n_jobs blocks (zav1_coeff_rate_batch)."""
    check(lib().zav1_coeff_rate_batch(tx_size, d_qcoeff.ptr, d_eob.ptr, d_jobs.ptr, n_jobs, d_costs.ptr,
                                      d_scans.ptr, d_rate.ptr, stream), 'coeff_rate_batch')


def rdoq_batch(tx_size, qindex, bd, d_tcoeff, d_qcoeff, d_dqcoeff, d_eob, d_jobs, n_jobs, d_costs, d_scans,
               cut_off_num=0, cut_off_denum=0, stream=None, d_work=None):
    """RDOQ of n_jobs blocks in place (zav1_rdoq_batch, or with a workspace
    DeviceArray zav1_rdoq_batch_ws)."""
    if d_work is None:
        check(lib().zav1_rdoq_batch(tx_size, qindex, bd, d_tcoeff.ptr, d_qcoeff.ptr, d_dqcoeff.ptr, d_eob.ptr,
                                    d_jobs.ptr, n_jobs, d_costs.ptr, d_scans.ptr, cut_off_num, cut_off_denum, stream),
              'rdoq_batch')
    else:
        check(lib().zav1_rdoq_batch_ws(tx_size, qindex, bd, d_tcoeff.ptr, d_qcoeff.ptr, d_dqcoeff.ptr, d_eob.ptr,
                                       d_jobs.ptr, n_jobs, d_costs.ptr, d_scans.ptr, cut_off_num, cut_off_denum,
                                       d_work.ptr, d_work.nbytes, stream), 'rdoq_batch_ws')


def sse_i32_batch(d_a, d_b, d_out, n_per_job, n_jobs, stream=None):
    check(lib().zav1_sse_i32_batch(d_a.ptr, d_b.ptr if d_b is not None else None, d_out.ptr, n_per_job,
                                   n_jobs, stream), 'sse_i32_batch')


def satd_batch(d_res, d_coeff, d_satd, n, n_jobs, stream=None):
    check(lib().zav1_satd_batch(d_res.ptr, d_coeff.ptr if d_coeff is not None else None, d_satd.ptr, n,
                                n_jobs, stream), 'satd_batch')


def scan_order(tx_size, scan_class):
    buf = np.zeros(1024, np.uint16)
    n = check(lib().zav1_scan_order(tx_size, scan_class, buf.ctypes.data), 'scan_order')
    return buf[:n].copy()


def txsearch_batch(d_plane, stride, bd, tx_size, d_xy, n_blocks, types, qindex, d_dist, d_eob, d_lsum,
                   stream=None):
    t = np.ascontiguousarray(types, dtype=np.uint8)
    check(lib().zav1_txsearch_batch(d_plane.ptr, stride, bd, tx_size, d_xy.ptr, n_blocks, t.ctypes.data, t.size,
                                    qindex, d_dist.ptr, d_eob.ptr, d_lsum.ptr, stream), 'txsearch_batch')


def txsearch_batch_q(d_plane, stride, bd, tx_size, d_xy, n_blocks, types, qindices, d_dist, d_eob, d_lsum,
                     stream=None):
    """The search at several qindices; outputs [q][block][type]."""
    t = np.ascontiguousarray(types, dtype=np.uint8)
    q = np.ascontiguousarray(qindices, dtype=np.int32)
    check(lib().zav1_txsearch_batch_q(d_plane.ptr, stride, bd, tx_size, d_xy.ptr, n_blocks, t.ctypes.data, t.size,
                                      q.ctypes.data, q.size, d_dist.ptr, d_eob.ptr, d_lsum.ptr, stream),
          'txsearch_batch_q')


def intra_pred_batch(d_plane, stride, bd, d_jobs, n_jobs, d_out, stream=None):
    check(lib().zav1_intra_pred_batch(d_plane.ptr, stride, bd, d_jobs.ptr, n_jobs, d_out.ptr, stream),
          'intra_pred_batch')


def cfl_batch(d_luma, stride, bd, d_jobs, n_jobs, d_dc, d_out, stream=None):
    check(lib().zav1_cfl_batch(d_luma.ptr, stride, bd, d_jobs.ptr, n_jobs, d_dc.ptr, d_out.ptr, stream), 'cfl_batch')


class DeblockGeom(ctypes.Structure):
    """zav1_deblock_geom: per-mi geometry, device pointers."""
    _fields_ = [('mi_cols', _i32), ('mi_rows', _i32), ('true_w', _i32), ('true_h', _i32), ('block_id', _vp),
                ('bw', _vp), ('bh', _vp), ('tw', _vp), ('th', _vp), ('skip', _vp)]


def deblock_geom(mi_cols, mi_rows, true_w, true_h, d):
    """d: dict of DeviceArrays block_id (u32), bw, bh, tw, th, optional skip (u8)."""
    return DeblockGeom(mi_cols, mi_rows, true_w, true_h, d['block_id'].ptr, d['bw'].ptr, d['bh'].ptr, d['tw'].ptr,
                       d['th'].ptr, d['skip'].ptr if 'skip' in d else None)


def deblock_plane(d_plane, stride, plane_w, plane_h, bd, plane, ss, level_v, level_h, sharpness, geom, stream=None):
    check(lib().zav1_deblock_plane(d_plane.ptr, stride, plane_w, plane_h, bd, plane, ss, level_v, level_h, sharpness,
                                   ctypes.byref(geom), stream), 'deblock_plane')


def cdef_search(rec, src, width, height, bd, qindex, subsampling, d_dirvar, d_mse_y, d_mse_uv, stream=None,
                d_skip=None, mi_cols=0, mi_rows=0):
    """rec/src: (y, u, v) DeviceArrays (u/v may be None); d_skip: optional
    per-mi signaled-skip map (u8, mi_rows x mi_cols)."""
    ptr = lambda a: a.ptr if a is not None else None  # noqa: E731
    check(lib().zav1_cdef_search(ptr(rec[0]), ptr(rec[1]), ptr(rec[2]), ptr(src[0]), ptr(src[1]), ptr(src[2]), width,
                                 height, bd, qindex, subsampling, ptr(d_skip), mi_cols, mi_rows, ptr(d_dirvar),
                                 d_mse_y.ptr, ptr(d_mse_uv), stream), 'cdef_search')


def cdef_apply(d_in, d_out, width, height, bd, damping, d_fb_strength, d_skip=None, mi_cols=0, mi_rows=0,
               d_stats=None, stream=None):
    """The CDEF frame pass (zav1_cdef_apply): d_in / d_out (y, u, v)
    DeviceArrays (u, v may be None), d_fb_strength (n_fb, 2) u8."""
    ptr = lambda a: a.ptr if a is not None else None  # noqa: E731
    check(lib().zav1_cdef_apply(ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]), ptr(d_out[0]), ptr(d_out[1]),
                                ptr(d_out[2]), width, height, bd, damping, d_fb_strength.ptr, ptr(d_skip), mi_cols,
                                mi_rows, ptr(d_stats), stream), 'cdef_apply')


WIENER_UNIT = np.dtype([('hfilter', '<i2', 8), ('vfilter', '<i2', 8), ('type', 'u1'), ('pad', 'u1', 7)])


def wiener_apply(d_dgd, d_out, plane_w, plane_h, bd, unit_size, ss_y, d_units, stream=None):
    check(lib().zav1_wiener_apply(d_dgd.ptr, d_out.ptr, plane_w, plane_h, bd, unit_size, ss_y, d_units.ptr, stream),
          'wiener_apply')


def wiener_refine(d_dgd, d_src, plane_w, plane_h, bd, win, unit_size, ss_y, d_units, d_err, use_refinement=True,
                  max_one_step=False, stream=None):
    check(lib().zav1_wiener_refine(d_dgd.ptr, d_src.ptr, plane_w, plane_h, bd, win, unit_size, ss_y,
                                   int(use_refinement), int(max_one_step), d_units.ptr, d_err.ptr, stream),
          'wiener_refine')


def rest_unit_grid(unit_size, plane_w, plane_h):
    nh, nv = _i32(), _i32()
    check(lib().zav1_rest_unit_grid(unit_size, plane_w, plane_h, ctypes.byref(nh), ctypes.byref(nv)), 'rest_unit_grid')
    return nh.value, nv.value


def wiener_stats(d_dgd, d_src, plane_w, plane_h, bd, win, unit_size, ss_y, d_M, d_H, stream=None):
    check(lib().zav1_wiener_stats(d_dgd.ptr, d_src.ptr, plane_w, plane_h, bd, win, unit_size, ss_y, d_M.ptr, d_H.ptr,
                                  stream), 'wiener_stats')


def unit_sse(d_a, d_b, plane_w, plane_h, bd, unit_size, ss_y, d_out, stream=None):
    check(lib().zav1_unit_sse(d_a.ptr, d_b.ptr, plane_w, plane_h, bd, unit_size, ss_y, d_out.ptr, stream), 'unit_sse')


def sgr_stats(d_dgd, d_src, plane_w, plane_h, bd, unit_size, ss_y, d_out, stream=None):
    check(lib().zav1_sgr_stats(d_dgd.ptr, d_src.ptr, plane_w, plane_h, bd, unit_size, ss_y, d_out.ptr, stream),
          'sgr_stats')


def txunit_batch(d_src, stride, plane_w, plane_h, bd, tx_size, d_pred, d_jobs, n_jobs, qindex, flags, d_dist, d_eob,
                 d_cul, d_qcoeff=None, d_recon=None, stream=None):
    """Fused transform units (zav1_txunit_batch); d_pred is a u16 DeviceArray."""
    ptr = lambda a: a.ptr if a is not None else None  # noqa: E731
    check(lib().zav1_txunit_batch(d_src.ptr, stride, plane_w, plane_h, bd, tx_size, d_pred.ptr,
                                  d_pred.nbytes // 2, d_jobs.ptr, n_jobs, qindex, flags, d_dist.ptr, d_eob.ptr,
                                  d_cul.ptr, ptr(d_qcoeff), ptr(d_recon), stream), 'txunit_batch')


def build_coeff_costs(base_qindex):
    """The frame's coefficient cost tables from the default CDFs (host, no
    device): int32[COEFF_COSTS_WORDS] in the zav1_coeff_costs layout."""
    out = np.zeros(COEFF_COSTS_WORDS, np.int32)
    check(lib().zav1_build_coeff_costs(int(base_qindex), out.ctypes.data), 'build_coeff_costs')
    return out


def intra_txt_rate(tx_size, intra_dir, tx_type):
    r = lib().zav1_intra_txt_rate(tx_size, intra_dir, tx_type)
    if r < 0:
        check(r, 'intra_txt_rate')
    return r


def txunit_rd_workspace(tx_size, n_jobs):
    b = ctypes.c_size_t()
    check(lib().zav1_txunit_rd_workspace(tx_size, n_jobs, ctypes.byref(b)), 'txunit_rd_workspace')
    return b.value


def txunit_rd_batch(d_src, stride, plane_w, plane_h, bd, tx_size, d_pred, d_jobs, n_jobs, qindex, flags, d_costs,
                    d_work, d_dist, d_eob, d_rate, d_cul, d_qcoeff=None, d_recon=None, cut_off=(0, 0),
                    stream=None):
    """Full-RD transform units (zav1_txunit_rd_batch); d_pred is u16."""
    ptr = lambda a: a.ptr if a is not None else None  # noqa: E731
    check(lib().zav1_txunit_rd_batch(d_src.ptr, stride, plane_w, plane_h, bd, tx_size, d_pred.ptr,
                                     d_pred.nbytes // 2, d_jobs.ptr, n_jobs, qindex, flags, cut_off[0], cut_off[1],
                                     d_costs.ptr, d_work.ptr, d_work.nbytes, d_dist.ptr, d_eob.ptr, d_rate.ptr,
                                     d_cul.ptr, ptr(d_qcoeff), ptr(d_recon), stream), 'txunit_rd_batch')


def pred_satd_batch(d_src, stride, plane_w, plane_h, bd, w, h, d_pred, d_jobs, n_jobs, d_satd, stream=None):
    """MDS0 SATD per (unit, prediction) job (zav1_pred_satd_batch); d_satd u32."""
    check(lib().zav1_pred_satd_batch(d_src.ptr, stride, plane_w, plane_h, bd, w, h, d_pred.ptr, d_pred.nbytes // 2,
                                     d_jobs.ptr, n_jobs, d_satd.ptr, stream), 'pred_satd_batch')


def md_fast_select(d_satd, d_rate, lam, n_units, n_cands, k, d_unit_xy, w, h, types, d_sel, d_sel_cost, d_jobs,
                   stream=None, xy_stride=2, survivors_only=False):
    """MDS0 survivors per unit and their transform-unit jobs (zav1_md_fast_select)."""
    tl = np.ascontiguousarray(types, dtype=np.uint8)
    check(lib().zav1_md_fast_select(d_satd.ptr, d_rate.ptr if d_rate is not None else None, int(lam), n_units,
                                    n_cands, k, d_unit_xy.ptr, xy_stride, w, h, tl.ctypes.data, len(tl),
                                    int(survivors_only), d_sel.ptr,
                                    d_sel_cost.ptr if d_sel_cost is not None else None, d_jobs.ptr, stream),
          'md_fast_select')


MD_UNIT = np.dtype([('x', '<i4'), ('y', '<i4'), ('n_top_px', '<i2'), ('n_topright_px', '<i2'), ('n_left_px', '<i2'),
                    ('n_bottomleft_px', '<i2'), ('edge_filter', 'u1'), ('filt_type', 'u1'), ('pad0', 'u1'),
                    ('pad1', 'u1')])
MD_CAND = np.dtype([('mode', 'u1'), ('delta', 'i1'), ('fi_mode', 'u1'), ('pad', 'u1')])


def md_fast_satd(d_src, d_recon, stride, plane_w, plane_h, bd, w, h, d_units, n_units, cands, d_satd, stream=None):
    """Fused MDS0 fast loop (zav1_md_fast_satd); cands: host MD_CAND array."""
    c = np.ascontiguousarray(cands, dtype=MD_CAND)
    check(lib().zav1_md_fast_satd(d_src.ptr, d_recon.ptr, stride, plane_w, plane_h, bd, w, h, d_units.ptr, n_units,
                                  c.ctypes.data, len(c), d_satd.ptr, stream), 'md_fast_satd')


class MdWave(ctypes.Structure):
    """zav1_md_wave (include/zav1_dsp.h)."""
    _fields_ = [('d_src', _vp), ('d_recon', _vp), ('stride', _i32), ('plane_w', _i32), ('plane_h', _i32),
                ('bd', _i32), ('w', _i32), ('h', _i32), ('d_units', _vp), ('n_units', _i32), ('cands', _vp),
                ('n_cands', _i32), ('d_rate', _vp), ('lam', ctypes.c_uint64), ('k', _i32), ('types', _vp),
                ('n_types', _i32), ('qindex', _i32), ('txu_flags', ctypes.c_uint), ('d_sel', _vp),
                ('d_sel_cost', _vp), ('d_dist', _vp), ('d_eob', _vp), ('d_cul', _vp), ('d_qcoeff', _vp),
                ('d_recon_out', _vp)]


class FrameSearch(ctypes.Structure):
    """zav1_frame_search: per tx size block / type counts, types and the
    device outputs of zav1_txsearch_frame."""
    _fields_ = [('n_blocks', ctypes.c_int32 * 19), ('n_types', ctypes.c_int32 * 19),
                ('types', (ctypes.c_uint8 * 16) * 19), ('dist', ctypes.c_void_p * 19), ('eob', ctypes.c_void_p * 19),
                ('lsum', ctypes.c_void_p * 19)]


class Context:
    """A device context: one stream and an arena of `budget` bytes."""

    def __init__(self, device=0, budget=1 << 30):
        h = _vp()
        check(lib().zav1_ctx_create(device, budget, ctypes.byref(h)), 'ctx_create')
        self.h = h.value

    def alloc(self, nbytes):
        p = _vp()
        check(lib().zav1_ctx_alloc(self.h, nbytes, ctypes.byref(p)), 'ctx_alloc')
        return p.value

    def usage(self):
        u, c = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().zav1_ctx_usage(self.h, ctypes.byref(u), ctypes.byref(c)), 'ctx_usage')
        return u.value, c.value

    def reset(self):
        check(lib().zav1_ctx_reset(self.h), 'ctx_reset')

    def stream(self):
        return lib().zav1_ctx_stream(self.h)

    def upload(self, plane, bd):
        a = np.ascontiguousarray(plane)
        p = _vp()
        check(lib().zav1_image_upload(self.h, a.ctypes.data, a.shape[1], a.shape[1], a.shape[0], bd,
                                      ctypes.byref(p)), 'image_upload')
        return p.value

    def txsearch_frame(self, d_plane, w, h, bd, qindices):
        """The tx-type search over every size of a plane (zav1_txsearch_frame);
        returns the FrameSearch record."""
        q = np.ascontiguousarray(qindices, dtype=np.int32)
        fs = FrameSearch()
        check(lib().zav1_txsearch_frame(self.h, d_plane, w, h, bd, q.ctypes.data, q.size, ctypes.byref(fs)),
              'txsearch_frame')
        return fs

    def submit_md_wave(self, wave):
        """Returns the status code (ZAV1_ENOMEM when the arena is too small)."""
        return lib().zav1_md_wave_submit(self.h, ctypes.byref(wave))

    def collect(self):
        check(lib().zav1_wave_collect(self.h), 'wave_collect')

    def close(self):
        if self.h:
            check(lib().zav1_ctx_destroy(self.h), 'ctx_destroy')
            self.h = None


def md_predict_survivors(d_recon, stride, plane_w, plane_h, bd, w, h, d_units, n_units, cands, d_sel, k, d_pred,
                         stream=None):
    c = np.ascontiguousarray(cands, dtype=MD_CAND)
    check(lib().zav1_md_predict_survivors(d_recon.ptr, stride, plane_w, plane_h, bd, w, h, d_units.ptr, n_units,
                                          c.ctypes.data, len(c), d_sel.ptr, k, d_pred.ptr, stream),
          'md_predict_survivors')


def deblock_level_sse(d_recon, d_src, plane_w, plane_h, bd, plane, ss, sharpness, geom, levels, d_work, d_sse,
                      stream=None):
    lv = np.ascontiguousarray(levels, dtype=np.uint8)
    check(lib().zav1_deblock_level_sse(d_recon.ptr, d_src.ptr, plane_w, plane_h, bd, plane, ss, sharpness,
                                       ctypes.byref(geom), lv.ctypes.data, lv.size, d_work.ptr, d_work.nbytes,
                                       d_sse.ptr, stream), 'deblock_level_sse')


def deblock_search_level(d_recon, d_src, plane_w, plane_h, bd, plane, ss, sharpness, geom, conv_th, d_work, d_sse,
                         stream=None):
    """search_filter_level on the GPU: (best level, sse at 0, best sse)."""
    b, e0, be = _i32(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().zav1_deblock_search_level(d_recon.ptr, d_src.ptr, plane_w, plane_h, bd, plane, ss, sharpness,
                                          ctypes.byref(geom), conv_th, d_work.ptr, d_work.nbytes, d_sse.ptr,
                                          ctypes.byref(b), ctypes.byref(e0), ctypes.byref(be), stream),
          'deblock_search_level')
    return b.value, e0.value, be.value
