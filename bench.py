"""Benchmark: the superblock-wavefront intra mode decision on one MI355X.

Workload (one step): a batch of N synthetic 4096x4096 10-bit luma planes
(1/f^2 spectral noise in [64, 940] + 30 hard lines, seeds 2000 + i) in flight
together, qp 40 (qindex 160), preset-4 funnel: 62 intra candidates per 16x16
leaf (13 modes, 8 x 6 angle deltas, FILTER_DC) -> prediction + Hadamard SATD
-> fast cost + exchange sort (15 survivors) -> DCT_DCT quantize_b units with
exact coefficient rate (MDS1) -> 4 survivors x the intra tx-type groups with
the rate and SATD gates -> RDOQ, exact rate, spatial distortion (MDS3) ->
full-cost winner, reconstruction committed for the next leaves.  Every SB
anti-diagonal step of every image runs as one chain of launches; the whole
batch is one HIP graph, replayed.  Default-CDF rate tables, key-frame
lambda.  Decisions and reconstruction are byte-identical to the oracle's
sequential restatement (tests) and to the plain per-call form (checked here
on a sample of leaves).

value = luma megapixels decided per second over the batch (N x 16.78 MP per
step; inputs resident in HBM).  Also reported: the plain per-call form on a
bounded sample of leaves (one launch per elementary call, selections on the
host), the per-stage device time of an eager pass (HIP events), the
roofline of the dominant kernel, and the oracle (CPU restatement) on a
bounded sample, multi-threaded over crops.

Extra stages (--extras): lf (loop-filter stage on the 4096x4096 8-bit
reconstruction), rdoq (the trellis over 262 144 16x16 candidates), txsearch
(exhaustive DC-predicted tx-type search, 2048x2048 10-bit), config4 (the
8192x4096 10-bit frame, SB128, 4x4 tiles, preset 2, through the
wavefront decision), serving (the config-5 mix through the wavefront
decision, footprint admission, p50 / p99 latency).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--images N] [--no-graph]
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'zenav1-svt_amd')

HBM_PEAK_GBS = 8000.0
# int32 VALU peak in the op-count model's units: 256 CU x 64 lanes x 2.4 GHz
# instructions, 2 ops each at best (the fused 3-operand forms the kernels
# issue, v_mad_*24 / v_add3 / v_lshl_add / v_max3, are 2 model ops, as an
# FMA is 2 flops).  The hardware view, VALU instructions issued per lane
# and clock, comes from the committed PMC counts (valu_issue_frac).
VALU_ISSUE_PEAK = 39.3e12
VALU_PEAK_TOPS = 78.6


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


COL_KIND = [0, 1, 0, 1, 2, 0, 2, 1, 2, 3, 0, 3, 1, 3, 2, 3]
ROW_KIND = [0, 0, 1, 1, 0, 2, 2, 2, 1, 3, 3, 0, 3, 1, 3, 2]
UD_FLIP = {4, 6, 8, 14}


def block_ops(ts_mod, tx_size, types, opsd, n_q=1):
    """Algorithmic integer ops of evaluating one block with every type in
    `types` at n_q quantizers: the DC prediction and residual once, each
    distinct column transform (kind, up-down flip) once, every type's row
    transforms and round shifts once, and per quantizer the fp quantize,
    distortion, eob and level sum.  1D transform counts are the
    straight-line op counts the generator records."""
    w, h = ts_mod.TX_DIMS[tx_size]

    def one(kind, n):
        return opsd[('dct' if kind == 0 else ('identity' if kind == 3 else 'adst')) + str(n)]
    n = w * h
    rect = 3 if (w == 2 * h or h == 2 * w) else 0
    ops = (w + h) + n * 2                       # DC prediction sums; residual + input shift
    variants = {(1 if COL_KIND[t] == 2 else COL_KIND[t], t in UD_FLIP) for t in types}
    ops += sum(w * one(k, h) + n * 2 for k, _ in variants)   # column passes + round shift
    for t in types:
        ops += h * one(ROW_KIND[t], w)
        ops += n * (2 + rect)                   # round shift (+ sqrt2 scale)
        ops += n_q * n * 12                     # fp quantize + dequant (abs, compare, add, mul, shift, sign x2)
        ops += n_q * n * 3                      # distortion (sub, mul, add)
        ops += n_q * n * 3                      # eob / level sum
    return ops


# Algorithmic integer ops per pixel of the self-guided statistics, per
# parameter set kind: box sums + A/B per position (r = 2 on every other row),
# the weighted A/B neighbourhood and output of each pass, and the five
# projection products.
SGR_OPS_R2 = (25 * 2 + 10) / 2 + 17
SGR_OPS_R1 = 9 * 2 + 10 + 25
SGR_OPS_PROJ = 13
SGR_OPS_PER_PX = (10 * (SGR_OPS_R2 + SGR_OPS_R1 + SGR_OPS_PROJ) + 4 * (SGR_OPS_R1 + SGR_OPS_PROJ) +
                  2 * (SGR_OPS_R2 + SGR_OPS_PROJ))


def pmc_kernel(kernel_prefix, bd=None, nq=None):
    """The committed PMC summary's name of a fused-search instance: the key
    starting with kernel_prefix whose template arguments carry the bit depth
    bd and end in nq quantizers per launch (the first match otherwise)."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
    if not os.path.exists(path):
        return kernel_prefix
    names = [k for k in json.load(open(path)).get('kernels', {}) if k.startswith(kernel_prefix)]
    exact = [k for k in names if (bd is None or ', %d, ' % bd in k) and (nq is None or k.endswith(', %d>' % nq))]
    return (exact or names or [kernel_prefix])[0]


def measured_valu_insts(kernel_prefix):
    """VALU instructions per launch of a kernel (SQ_INSTS_VALU, wave-level,
    from the committed PMC summary), or None."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
    if not os.path.exists(path):
        return None
    for name, k in json.load(open(path)).get('kernels', {}).items():
        if name.startswith(kernel_prefix) and 'valu_insts' in k:
            return k['valu_insts']
    return None


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of a kernel from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor the
    same counter shows on a device-to-device copy of known size), or None."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    for name, k in d.get('kernels', {}).items():
        if name.startswith(kernel_prefix):
            return int(k['fetch_kib'] * 1024 * d['fetch_scale'] + k['write_kib'] * 1024 * d['write_scale'])
    return None


def _short_kernel_name(name):
    """'void zav1::k_rdoq_wave<2>(int const*, ...)' -> 'k_rdoq_wave<2>'."""
    for ns in ('(anonymous namespace)::', 'zav1::'):
        name = name.replace(ns, '')
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == '<':
            depth += 1
        elif ch == '>':
            depth -= 1
        elif ch == '(' and depth == 0:
            cut = i
            break
    n = name[:cut]
    if n.startswith('void '):
        n = n[5:]
    return n.strip()


def pmc_summary(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, merge_into=None):
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from
    rocprofv3 --pmc CSV output directories, with the byte scale of each
    counter from the 64 MiB device copy runs (--copy-calibration); merged
    over the kernels of an existing summary when given."""
    import csv
    import glob

    def per_kernel(d, counter):
        acc = {}
        for f in glob.glob(os.path.join(d, '**', '*counter_collection.csv'), recursive=True):
            for row in csv.DictReader(open(f)):
                if row.get('Counter_Name') != counter:
                    continue
                k = _short_kernel_name(row['Kernel_Name'])
                a = acc.setdefault(k, [0.0, 0, 0])
                a[0] += float(row['Counter_Value'])
                a[1] += 1
                a[2] = max(a[2], int(float(row.get('Grid_Size', 0) or 0)))
        return {k: (v[0] / v[1], v[1], v[2]) for k, v in acc.items()}

    fe, wr = per_kernel(fetch_dir, 'FETCH_SIZE'), per_kernel(write_dir, 'WRITE_SIZE')
    cf, cw = per_kernel(cal_fetch_dir, 'FETCH_SIZE'), per_kernel(cal_write_dir, 'WRITE_SIZE')
    copy = [k for k in cf if 'copyBuffer' in k]
    if not copy:
        raise SystemExit('no copy kernel in the calibration runs')
    n_kib = (64 << 20) / 1024.0
    out = json.load(open(merge_into)) if merge_into else {'kernels': {}}
    out['fetch_scale'] = round(n_kib / cf[copy[0]][0], 4)
    out['write_scale'] = round(n_kib / cw[[k for k in cw if 'copyBuffer' in k][0]][0], 4)
    out['calibration'] = '__amd_rocclr_copyBuffer of 64 MiB: FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB per dispatch' % (
        cf[copy[0]][0], cw[[k for k in cw if 'copyBuffer' in k][0]][0])
    for k in fe:
        if k in wr and 'copyBuffer' not in k:
            e = out['kernels'].get(k, {})
            e.update({'fetch_kib': round(fe[k][0], 1), 'write_kib': round(wr[k][0], 1), 'dispatches': fe[k][1],
                      'grid_size': fe[k][2], 'per_dispatch': 'mean'})
            out['kernels'][k] = e
    return out


def copy_calibration(zav1):
    """Device-to-device copies of 64 MiB for the PMC byte-counter scale."""
    n = 64 << 20
    a, b = zav1.DeviceArray(n, np.uint8), zav1.DeviceArray(n, np.uint8)
    a.zero()
    for _ in range(3):
        zav1.check(zav1.lib().zav1_memcpy_d2d(b.ptr, a.ptr, n, None), 'memcpy_d2d')
    zav1.sync(None)
    print(json.dumps({'copy_calibration_bytes': n, 'copies': 3}))


def loopfilter_stage(zav1, synth, args, cpu=True):
    L = zav1.lib()
    lfm = _load('zav1_loopfilter', os.path.join(PKG, 'loopfilter.py'))
    N = 4096
    src_y, rec_y = synth.config3_recon(N, N, seed=3000)
    (su, ru), (sv, rv) = synth.config3_chroma(N // 2, N // 2, seed=3000)
    geom = synth.partition_geometry(N // 4, N // 4, 3001, skip_frac=0.2)
    qi = synth.qp_to_qindex(32)
    levels = synth.key_frame_lf_levels(int(zav1.quant_row(qi, 8).dequant[1]), 8)
    st = lfm.LoopFilterStage(zav1, {'src': (src_y, su, sv), 'rec': (rec_y, ru, rv)}, geom, 8, qi, levels)
    sp = ctypes.c_void_p()
    zav1.check(L.zav1_stream_create(ctypes.byref(sp)), 'stream_create')
    stream = sp.value
    for _ in range(2):
        st.run(stream)
    zav1.sync(stream)
    nph = len(st.PHASES)
    ev = []
    for _ in range(nph + 1):
        e = ctypes.c_void_p()
        L.zav1_event_create(ctypes.byref(e))
        ev.append(e.value)
    # luma self-guided launch alone, for the kernel roofline
    e_sg = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in e_sg:
        L.zav1_event_create(ctypes.byref(e))
    phase_ms = np.zeros(nph)
    sg_luma_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.lf_steps):
        L.zav1_event_record(ev[0], stream)
        for i, ph in enumerate(st.PHASES):
            if ph == 'selfguided':
                L.zav1_event_record(e_sg[0].value, stream)
                zav1.sgr_stats(st.dgd[0], st.src[0], st.w, st.h, 8, st.ru[0], 0, st.sgrs[0], stream)
                L.zav1_event_record(e_sg[1].value, stream)
                for p in (1, 2):
                    zav1.sgr_stats(st.dgd[p], st.src[p], st.cw, st.ch, 8, st.ru[p], 1, st.sgrs[p], stream)
            else:
                getattr(st, ph)(stream)
            L.zav1_event_record(ev[i + 1], stream)
        zav1.sync(stream)
        for i in range(nph):
            ms = ctypes.c_float()
            L.zav1_event_elapsed_ms(ev[i], ev[i + 1], ctypes.byref(ms))
            phase_ms[i] += ms.value
        ms = ctypes.c_float()
        L.zav1_event_elapsed_ms(e_sg[0].value, e_sg[1].value, ctypes.byref(ms))
        sg_luma_ms += ms.value
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000.0 / args.lf_steps
    sg_s = sg_luma_ms / 1000.0 / args.lf_steps
    sg_ops = N * N * SGR_OPS_PER_PX
    out = {'workload': 'config3 loop-filter stage: 4096x4096 8-bit 4:2:0, deblock + CDEF 64-strength SSE tables '
                       '(20% of blocks signaled skip) + Wiener M/H + unit SSE + self-guided stats '
                       '(RU 128 luma / 64 chroma)',
           'value': round(N * N / 1e6 / (ms_step / 1000.0), 2), 'unit': 'MP/s', 'steps': args.lf_steps,
           'ms_per_step': round(ms_step, 3),
           'phase_ms': {ph: round(phase_ms[i] / args.lf_steps, 4) for i, ph in enumerate(st.PHASES)},
           'roofline': {'bound': 'valu', 'kernel': 'k_sgr_strip<uint8_t> (luma)', 'unit': 'Tops/s',
                        'achieved': round(sg_ops / sg_s / 1e12, 3), 'peak': VALU_PEAK_TOPS,
                        'frac': round(sg_ops / sg_s / 1e12 / VALU_PEAK_TOPS, 4),
                        'traffic': measured_traffic('k_sgr_strip'),
                        'avg_launch_us': round(sg_s * 1e6, 2), 'algorithmic_ops_per_launch': int(sg_ops),
                        'hbm_algorithmic_GBs': round(2 * N * N / sg_s / 1e9, 2)}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        c = 2048
        ys, yr = np.ascontiguousarray(src_y[:c, :c]), np.ascontiguousarray(rec_y[:c, :c])
        us_, ur_ = np.ascontiguousarray(su[:c // 2, :c // 2]), np.ascontiguousarray(ru[:c // 2, :c // 2])
        vs_, vr_ = np.ascontiguousarray(sv[:c // 2, :c // 2]), np.ascontiguousarray(rv[:c // 2, :c // 2])
        g = {k: np.ascontiguousarray(v[:c // 4, :c // 4]) for k, v in geom.items() if k != 'skip'}
        gskip = np.ascontiguousarray(geom['skip'][:c // 4, :c // 4])
        LR = O.lrlib()
        tc0 = time.perf_counter()
        d = [O.deblock_plane(yr, 8, 0, 0, levels[0], levels[1], 0, g, c // 4, c // 4, c, c),
             O.deblock_plane(ur_, 8, 1, 1, levels[2], levels[2], 0, g, c // 4, c // 4, c, c),
             O.deblock_plane(vr_, 8, 2, 1, levels[3], levels[3], 0, g, c // 4, c // 4, c, c)]
        for fbr in range(c // 64):
            for fbc in range(c // 64):
                O.cdef_search_fb(tuple(d), (ys, us_, vs_), c, c, 8, qi, 1, fbr, fbc, np.arange(64), skip=gskip)
        for p, (pl, sp_, us, ss) in enumerate([(d[0], ys, 128, 0), (d[1], us_, 64, 1), (d[2], vs_, 64, 1)]):
            pw = pl.shape[1]
            n = LR.or_count_units(us, pw)
            for uy in range(n):
                for ux in range(n):
                    lim = O.unit_limits(us, pw, pw, ss, ux, uy)
                    O.wiener_stats(7 if p == 0 else 5, pl, sp_, 8, lim)
                    LR.or_sse_region(pl.ctypes.data, sp_.ctypes.data, pw, 8, *lim)
                    for s_ in range(16):
                        O.sgr_unit_stats(pl, sp_, 8, lim, s_)
        tc1 = time.perf_counter()
        out['cpu_baseline'] = {'value': round(c * c / 1e6 / (tc1 - tc0), 4), 'unit': 'MP/s', 'cores': 1,
                               'kind': 'port', 'sample': '2048x2048 crop of the same frame, whole stage '
                                                         '(oracle C, one thread)',
                               'seconds': round(tc1 - tc0, 2)}
    return out


# Per-sample int ops of one candidate in the fused fast loop, by prediction
# class (counted from the predictor arithmetic): DC / V / H copy 1, SMOOTH 10,
# SMOOTH_V / _H 5, PAETH 10, directional interpolation 9, filter intra 16;
# plus the residual (1) and the Hadamard + absolute sum (8 per sample for
# 4x4 / 8x8 tiles, +3 with the 16x16 combine, +6 with the 32x32 one).
def md_fast_ops_per_unit(cands, w, h):
    t = min(w, h, 32)
    had = 8 + (3 if t >= 16 else 0) + (3 if t >= 32 else 0)
    tot = 0
    for mode, delta, fi in cands:
        if fi < 5:
            p = 16
        elif mode in (3, 4, 5, 6, 7, 8) or (mode in (1, 2) and delta != 0):
            p = 9
        elif mode == 9 or mode == 12:
            p = 10
        elif mode in (10, 11):
            p = 5
        else:
            p = 1
        tot += (p + 1 + had) * w * h
    return tot


def rdoq_stage(zav1, synth, args, cpu=True):
    """RDOQ (zav1_rdoq_batch) of every tx-type candidate of every 16x16 block
    of the config-2 frame (2048x2048 10-bit): 16 384 blocks x 16 types =
    262 144 trellis runs, as the tx-type search asks for them.  Residual =
    block minus its rounded mean; each type's forward transform on the
    device, fp quantization at qp 20, eob from the type's scan (all
    untimed); the timed launch requantizes every candidate in place (the
    quantized inputs restored by a device copy before each step, outside the
    events).  Rate tables: synthetic (synth.coeff_costs), rdmult 1054880."""
    L = zav1.lib()
    N, ts, n, nt = 2048, 2, 256, 16
    img = synth.config2_luma(N, N, seed=2000).astype(np.int32)
    blk = img.reshape(N // 16, 16, N // 16, 16).transpose(0, 2, 1, 3).reshape(-1, n)
    res = (blk - (blk.sum(axis=1, keepdims=True) + n // 2) // n).astype(np.int32)
    nb = res.shape[0]
    nj = nb * nt
    types = np.tile(np.arange(nt, dtype=np.uint8), nb)  # job j: block j // 16, type j % 16
    qi = synth.qp_to_qindex(20)
    D = zav1.DeviceArray
    d_res = D.from_host(np.repeat(res, nt, axis=0))
    d_coef = D((nj, n), np.int32)
    zav1.fwd_txfm2d_batch(d_res, d_coef, D.from_host(types), ts, nj)
    del d_res
    d_q, d_dq = D((nj, n), np.int32), D((nj, n), np.int32)
    zav1.quantize_batch(d_coef, d_q, d_dq, D.from_host(np.full(nj, qi, np.uint8)), n, nj, 0, 10, 0)
    scans = np.concatenate([zav1.scan_order(ts, c) for c in (0, 1, 2)]).astype(np.uint16)
    q0, dq0 = d_q.to_host(), d_dq.to_host()
    # tx class of each type (V_* vertical, H_* horizontal) and its scan
    cls_of = np.array([0] * 10 + [2, 1, 2, 1, 2, 1], np.uint8)
    scan_of = {0: scans[:n], 2: scans[n:2 * n], 1: scans[2 * n:]}
    eob0 = np.zeros(nj, np.uint16)
    for c in (0, 1, 2):
        sel = np.nonzero(cls_of[types] == c)[0]
        nzq = q0[sel][:, scan_of[c]] != 0
        eob0[sel] = np.where(nzq.any(axis=1), n - np.argmax(nzq[:, ::-1], axis=1), 0)
    jobs = np.zeros(nj, zav1.RDOQ_JOB)
    jobs['rdmult'] = 1054880
    jobs['coeff'] = np.arange(nj, dtype=np.uint32) * n
    jobs['tx_class'] = cls_of[types]
    costs = synth.coeff_costs(2020)
    d_jobs, d_costs, d_scans = D.from_host(jobs), D.from_host(costs), D.from_host(scans)
    d_eob, d_q_keep, d_dq_keep, d_eob_keep = D.from_host(eob0), D.from_host(q0), D.from_host(dq0), D.from_host(eob0)
    st = ctypes.c_void_p()
    zav1.check(L.zav1_stream_create(ctypes.byref(st)), 'stream_create')
    stream = st.value
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        L.zav1_event_create(ctypes.byref(e))

    def restore():
        for dst, src in ((d_q, d_q_keep), (d_dq, d_dq_keep), (d_eob, d_eob_keep)):
            zav1.check(L.zav1_memcpy_d2d(dst.ptr, src.ptr, src.nbytes, stream), 'memcpy_d2d')

    d_ws = D((L.zav1_rdoq_workspace(ts, nj),), np.uint8)
    use_ws = [True]

    def launch():
        zav1.rdoq_batch(ts, qi, 10, d_coef, d_q, d_dq, d_eob, d_jobs, nj, d_costs, d_scans, stream=stream,
                        d_work=d_ws if use_ws[0] else None)

    steps = max(5, args.steps // 5)

    def time_form(lane):
        # the lane form (a workspace, threshold 0) or the latency form (none: the
        # group form at this size)
        use_ws[0] = lane
        prev = L.zav1_rdoq_lane_min(ts, 0)
        try:
            for _ in range(2):
                restore()
                launch()
            zav1.sync(stream)
            tot = 0.0
            for _ in range(steps):
                restore()
                L.zav1_event_record(ev[0].value, stream)
                launch()
                L.zav1_event_record(ev[1].value, stream)
                zav1.sync(stream)
                ms = ctypes.c_float()
                L.zav1_event_elapsed_ms(ev[0].value, ev[1].value, ctypes.byref(ms))
                tot += ms.value
            return tot / steps, d_q.to_host()
        finally:
            L.zav1_rdoq_lane_min(ts, prev)

    group_ms, q_group = time_form(False)
    lane_ms, q_lane = time_form(True)
    assert np.array_equal(q_group, q_lane), 'trellis forms disagree'
    form = 'group' if group_ms < lane_ms else 'lane'
    ms_step = min(group_ms, lane_ms)
    # coefficient rate of every candidate after the trellis (one launch)
    rjobs = np.zeros(nj, zav1.RATE_JOB)
    rjobs['coeff'] = jobs['coeff']
    rjobs['tx_class'] = jobs['tx_class']
    d_rjobs, d_rate = D.from_host(rjobs), D((nj,), np.int32)
    rate_ms = 0.0
    for it in range(steps + 1):
        L.zav1_event_record(ev[0].value, stream)
        zav1.coeff_rate_batch(ts, d_q, d_eob, d_rjobs, nj, d_costs, d_scans, d_rate, stream=stream)
        L.zav1_event_record(ev[1].value, stream)
        zav1.sync(stream)
        ms = ctypes.c_float()
        L.zav1_event_elapsed_ms(ev[0].value, ev[1].value, ctypes.byref(ms))
        rate_ms += ms.value if it else 0.0
    rate_ms /= steps
    eob1 = d_eob.to_host()
    q1 = d_q.to_host()
    steps_coef = int(eob0.astype(np.int64).sum())
    # algorithmic HBM bytes: tcoeff, qcoeff, dqcoeff read, changed levels written back, jobs and eobs
    changed = int((q1 != q0).sum())
    alg_bytes = nj * n * 12 + changed * 8 + nj * (24 + 4)
    out = {'workload': 'RDOQ of every tx-type candidate (16) of the 16 384 16x16 blocks of the 2048x2048 10-bit '
                       'config-2 frame at qp 20 (block-mean residual, fp quantization; synthetic rate tables)',
           'value': round(N * N / 1e6 / (ms_step / 1000.0), 2), 'unit': 'MP/s (all 16 candidates per pixel)',
           'steps': steps, 'ms_per_step': round(ms_step, 4), 'trellis_runs': nj,
           'trellis_runs_per_s': round(nj / (ms_step / 1000.0)),
           'trellis_steps_per_s': round(steps_coef / (ms_step / 1000.0)),
           'mean_eob_in': round(float(eob0.mean()), 2), 'mean_eob_out': round(float(eob1.mean()), 2),
           'levels_changed_frac': round(changed / max(1, int((q0 != 0).sum())), 4),
           'coeff_rate_ms': round(rate_ms, 4), 'coeff_rate_blocks_per_s': round(nj / (rate_ms / 1000.0)),
           'lane_form_ms': round(lane_ms, 4), 'group_form_ms': round(group_ms, 4),
           'roofline': {'bound': 'latency', 'kernel': ('k_rdoq_group<TX_16X16, 16>' if form == 'group' else
                                                       'k_rdoq<TX_16X16>'), 'unit': 'GB/s',
                        'achieved': round(alg_bytes / (ms_step / 1000.0) / 1e9, 2), 'peak': 8000.0,
                        'frac': round(alg_bytes / (ms_step / 1000.0) / 1e9 / 8000.0, 4),
                        'traffic': measured_traffic('k_rdoq<2>'), 'avg_launch_us': round(ms_step * 1000.0, 2),
                        'algorithmic_bytes_per_launch': alg_bytes}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        k = 16384
        coef = d_coef.to_host()
        row = O.quant_row(qi, 10)
        dqv = [row.dequant[0], row.dequant[1]]
        tc, ec = synth.tables_for(costs, 2, 4, 0)
        tc0 = time.perf_counter()
        for j in range(k):
            if eob0[j]:
                c = int(cls_of[types[j]])
                O.optimize_b(coef[j], q0[j], dq0[j], int(eob0[j]), scan_of[c], dqv, tc, ec, 1054880, ts, c)
        tc1 = time.perf_counter()
        out['cpu_baseline'] = {'value': round(k / nt * n / 1e6 / (tc1 - tc0), 4),
                               'unit': 'MP/s (all 16 candidates per pixel)', 'cores': 1, 'kind': 'port',
                               'sample': 'the first %d candidates (1024 blocks x 16 types; oracle C optimize_b, '
                                         'one thread, ctypes call per block)' % k,
                               'seconds': round(tc1 - tc0, 3)}
    return out


def config4_stage(zav1, synth, args):
    """Config 4: the 8192x4096 10-bit frame (four seeded 2048-wide tile
    columns of 1/f noise with 25% screen regions), SB128, 4x4 tiles,
    preset 2, qp 28, through the wavefront mode decision with all 16
    tiles' wavefronts in the same launches (one HIP graph), timed over
    graph replays; graph == eager checked."""
    wf = _load('zav1_wavefront', os.path.join(PKG, 'wavefront.py'))
    L = zav1.lib()
    W, H, qp, preset = 8192, 4096, 28, 2
    img = np.concatenate([synth.config4_tile_column(H, 2048, seed=4000 + k) for k in range(4)], axis=1)
    qidx = synth.qp_to_qindex(qp)
    lam = int(L.zav1_kf_full_lambda(qidx, qp, 10))
    cfg = wf.funnel_cfg(W, H, 10, qidx, lam, preset, qp, sb=128, tiles=(4, 4))
    md = wf.WavefrontMD(zav1, cfg)
    md.upload(img[None])
    st = zav1.stream_create()
    md.run(True, st)
    zav1.stream_sync(st)
    tuned = md.autotune(stream=st)
    md.run(True, st)
    zav1.stream_sync(st)
    steps = max(2, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        md.run(True, st)
    zav1.stream_sync(st)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    rec, dec = md.recon(), md.raw_decisions()
    md.run(False, st)
    zav1.stream_sync(st)
    same = bool(np.array_equal(md.recon(), rec) and np.array_equal(md.raw_decisions(), dec))
    info = md.info
    md.close()
    # the frame encoded: the same decision with the levels kept, then its
    # stream (16 tiles on 16 host threads) -- one frame, not pipelined
    cfg[0]['keep_levels'] = 1
    mde = wf.WavefrontMD(zav1, cfg)
    mde.upload(img[None])
    mde.run(True, st)
    mde.autotune(stream=st)
    mde.run(True, st)
    first = mde.write_stills(0, 1, threads=16, stream=st)[0]
    enc_ms, stream_same = [], True
    for _ in range(steps):
        t0 = time.perf_counter()
        mde.run(True, st)
        got = mde.write_stills(0, 1, threads=16, stream=st)[0]
        enc_ms.append((time.perf_counter() - t0) * 1000.0)
        stream_same &= got == first
    mde.close()
    zav1.stream_destroy(st)
    enc = float(np.median(enc_ms))
    return {'metric': 'MP/s through the wavefront mode decision, 16 tiles co-scheduled',
            'value': round(W * H / 1e6 / (ms / 1000.0), 2), 'unit': 'MP/s', 'ms_per_step': round(ms, 2),
            'steps': steps, 'graph_vs_eager_identical': same, 'wavefront_steps': info.n_steps, 'autotune': tuned,
            'encoded': {'value': round(W * H / 1e6 / (enc / 1000.0), 2), 'unit': 'MP/s', 'ms_per_frame': round(enc, 2),
                        'stream_bytes': len(first), 'streams_identical': bool(stream_same),
                        'what': 'decision + AV1 stream of the frame (16 tiles on 16 host threads), one frame at a '
                                'time, median of the timed frames'},
            'max_leaves_per_step': info.max_leaves_per_step,
            'config': {'workload': 'config4: 8192x4096 10-bit luma (1/f noise + 25%% screen regions), SB128, 4x4 '
                                   'tiles, preset 2 (%d candidates, %d -> %d survivors), qp 28, 16x16 leaves'
                                   % (int(cfg[0]['n_cands']), int(cfg[0]['k1']), int(cfg[0]['k3']))}}


def preset8_stage(zav1, synth, args):
    """The luma path at preset 8 end to end, as the reference runs it there
    (the partition pre-pass decides the tree, the mode decision codes its
    leaves, pipeline.rs:8000-8060): per batch of images, the PD0 pass
    (level 1, variance cap, small-eob rate) on the device, its leaf maps
    read back, a wavefront plan made from them on the host (leaves 8x8 ..
    64x64, their coding-order dependencies), the preset-8 funnel over the
    leaves on the device (eager: a plan serves one batch), each image's
    stream written on host threads.  Timed once part by part, then
    pipelined: every batch's plan, decision, streams and release on a
    worker thread with its own stream, up to three batches in flight (their
    small steps overlap on the device), the PD0 passes on the calling
    thread.  Every pipelined batch's streams must equal the part-by-part
    batch's."""
    from concurrent.futures import ThreadPoolExecutor
    wf = _load('zav1_wavefront', os.path.join(PKG, 'wavefront.py'))
    part = _load('zav1_partition', os.path.join(PKG, 'partition.py'))
    bsm = _load('zav1_bitstream', os.path.join(PKG, 'bitstream.py'))
    L = zav1.lib()
    S, n, qp, preset, bd = args.size, args.p8_images, 40, 8, 10
    qidx = synth.qp_to_qindex(qp)
    lam = int(L.zav1_kf_full_lambda(qidx, qp, bd))
    cores = min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(cores) as ex:
        imgs = np.stack(list(ex.map(lambda i: synth.config2_luma(S, S, seed=2000 + i), range(n))))
    tc = bsm.legal_tile_cols(S, S, 64)
    tiles = (tc, 1)
    # the pass's tiles are given in 64x64 units: the frame's tile columns
    pcfg = part.pd0_config(zav1, S, S, bd, qp, qidx, preset, n_images=n, tiles=(-(-(S // 64) // tc), S // 64))
    d_plane = zav1.DeviceArray.from_host(np.ascontiguousarray(imgs.reshape(n * S, S), np.uint16))
    pp = part.PartitionPass(zav1, pcfg)
    st_p = zav1.stream_create()
    sts = [zav1.stream_create() for _ in range(2)]

    def pd0():
        pp.run(d_plane, st_p)
        return np.ascontiguousarray(pp.d_leaf.to_host(st_p).reshape(-1))  # synchronous read-back

    def make_plan(lm, st):
        cfg = wf.tree_cfg(S, S, bd, qidx, lam, preset, qp, lm, n_images=n, tiles=tiles)
        cfg[0]['keep_levels'] = 1
        md = wf.WavefrontMD(zav1, cfg)
        zav1.check(L.zav1_memcpy_d2d(md.info.d_src, d_plane.ptr, d_plane.nbytes, st), 'memcpy_d2d')
        return md

    def write_and_release(md, st):
        got = md.write_stills(threads=cores, stream=st)
        md.close()
        return got

    if args.p8_profile_pass:  # one batch and nothing else, for profilers
        md = make_plan(pd0(), sts[0])
        md.run(False, sts[0])
        got = write_and_release(md, sts[0])
        pp.close()
        d_plane.free()
        for s_ in [st_p] + sts:
            zav1.stream_destroy(s_)
        return {'profile_pass': 'preset 8, %d x %dx%d 10-bit: PD0, plan, eager decision, streams' % (n, S, S),
                'stream_bytes': sum(len(x) for x in got)}
    # part by part, one batch (after one untimed batch)
    md = make_plan(pd0(), sts[0])
    md.run(False, sts[0])
    write_and_release(md, sts[0])
    part_ms = {}
    t = time.perf_counter()
    lm = pd0()
    part_ms['pd0_and_map_readback'] = time.perf_counter() - t
    t = time.perf_counter()
    md = make_plan(lm, sts[0])
    zav1.stream_sync(sts[0])
    part_ms['plan'] = time.perf_counter() - t
    info = md.info
    n_leaves, n_steps = info.n_leaves, info.n_steps
    t = time.perf_counter()
    md.run(False, sts[0])
    zav1.stream_sync(sts[0])
    part_ms['decision'] = time.perf_counter() - t
    t = time.perf_counter()
    ref = md.write_stills(threads=cores, stream=sts[0])
    part_ms['streams'] = time.perf_counter() - t
    t = time.perf_counter()
    md.close()
    part_ms['release'] = time.perf_counter() - t
    seq_ms = sum(part_ms.values()) * 1000.0
    cnt = np.bincount(lm, minlength=7)
    leaves_by_size = {'%dx%d' % (1 << v, 1 << v): int(cnt[v] // ((1 << v) // 4) ** 2) for v in range(3, 7)}
    # pipelined: each batch's plan, decision and streams on a worker thread
    # of its own (its own stream), up to `depth` batches in flight; the PD0
    # passes (shared buffers) stay on this thread, one batch ahead
    import threading
    K, depth = max(2, args.steps), 3
    local = threading.local()
    wsts = []

    def batch(lm):
        if not hasattr(local, 'st'):
            local.st = zav1.stream_create()
            wsts.append(local.st)
        m = make_plan(lm, local.st)
        m.run(False, local.st)
        return write_and_release(m, local.st)

    same = True
    with ThreadPoolExecutor(depth) as pool:
        t0 = time.perf_counter()
        futs = []
        for k in range(K):
            if len(futs) >= depth:
                same &= futs[len(futs) - depth].result() == ref
            futs.append(pool.submit(batch, pd0()))
        for f in futs[max(0, len(futs) - depth):]:
            same &= f.result() == ref
        t1 = time.perf_counter()
    for s_ in wsts:
        zav1.stream_destroy(s_)
    pp.close()
    d_plane.free()
    for s in [st_p] + sts:
        zav1.stream_destroy(s)
    ms = (t1 - t0) * 1000.0 / K
    mp = n * S * S / 1e6
    nbytes = sum(len(x) for x in ref)
    return {'metric': 'MP/s encoded at preset 8 (PD0 tree + mode decision of its leaves + streams)',
            'value': round(mp / (ms / 1000.0), 2), 'unit': 'MP/s', 'ms_per_batch': round(ms, 2), 'batches': K,
            'part_by_part': {'value': round(mp / (seq_ms / 1000.0), 2), 'ms_per_batch': round(seq_ms, 2),
                             'ms': {k: round(v * 1000.0, 2) for k, v in part_ms.items()}},
            'streams_identical': bool(same), 'stream_bytes_per_batch': nbytes,
            'bits_per_pixel': round(8.0 * nbytes / (n * S * S), 4),
            'leaves': n_leaves, 'leaves_by_size': leaves_by_size, 'wavefront_steps': n_steps, 'host_threads': cores,
            'config': {'workload': 'preset 8: %d x %dx%d 10-bit luma (the headline content), qp 40, PD0 level 1 '
                                   'tree, preset-8 funnel over its leaves, SB64, %dx%d tiles' % (n, S, S, tc, 1)}}


def serving_stage(zav1, synth, args):
    """Serving mix (config 5) encoded at a fixed Poisson arrival rate:
    requests of log-uniform area 256^2..4096x3072, 8 / 10 bits, qp 20..50,
    preset 4 / 6 / 8, admitted against an HBM budget, continuous batching at
    wavefront-step granularity, each request's stream written on a host pool
    as its decision completes.  The rate is args.serve_load of the loop's
    measured back-to-back throughput (all requests at once)."""
    wf = _load('zav1_wavefront', os.path.join(PKG, 'wavefront.py'))
    srv = _load('zav1_serving', os.path.join(PKG, 'serving.py'))
    reqs = synth.serving_mix(args.serve_n, seed=5000)
    imgs = synth.serving_images(reqs, seed=5000)
    specs = [(img, r[2], r[3], r[4]) for img, r in zip(imgs, reqs)]
    budget = args.serve_budget_gb << 30
    server = srv.WavefrontServer(zav1, wf, synth, budget, max_active=args.serve_active, chunk=args.serve_chunk,
                                 encode=True, writers=4, write_threads=4)
    _, st0 = server.run(specs, [0.0] * len(specs), keep_results=False)  # back to back
    mp = st0['megapixels']
    b2b = mp / st0['wall_s']
    rate = args.serve_load * b2b / (mp / len(specs))
    rng = np.random.default_rng(5002)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(specs)))
    _, st = server.run(specs, arrivals, keep_results=False)
    server.close()
    lat = st['latency_s']
    return {'metric': 'per-image encode latency (ms): arrival to AV1 stream, Poisson arrivals',
            'p50_ms': round(float(np.percentile(lat, 50)) * 1e3, 2),
            'p99_ms': round(float(np.percentile(lat, 99)) * 1e3, 2),
            'mean_ms': round(float(lat.mean()) * 1e3, 2),
            'value': round(st['megapixels'] / st['wall_s'], 2), 'unit': 'MP/s', 'requests': len(specs),
            'megapixels': round(st['megapixels'], 2), 'arrival_rate_hz': round(rate, 2), 'load': args.serve_load,
            'back_to_back_MPs': round(b2b, 2), 'stream_bytes': int(st['stream_bytes']),
            'admission': {'hbm_budget_GB': args.serve_budget_gb, 'max_active': args.serve_active,
                          'chunk_steps': args.serve_chunk, 'peak_active': st['peak_active'],
                          'peak_MB': round(st['peak_bytes'] / 1e6, 1),
                          'deferred_by_memory': st['deferred_by_memory']},
            'config': {'workload': 'serving mix: area log-uniform 256^2..4096x3072, aspect 1:1..16:9, 50/50 '
                                   '8/10-bit, qp U{20..50}, preset U{4,6,8}; per request upload, the wavefront '
                                   'mode decision (16x16 leaves, the fewest tile columns AV1 allows) and its AV1 '
                                   'still stream written on a host pool (4 writers x 4 tile threads)'}}


def cpu_info():
    """Host CPU model and the threads this process may run on."""
    model = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return model, sorted(os.sched_getaffinity(0))


def oracle_cflags():
    try:
        for line in open(os.path.join(ROOT, 'oracle', 'Makefile')):
            if line.startswith('CFLAGS'):
                return line.split('=', 1)[1].strip()
    except OSError:
        pass
    return None


def wavefront_cpu_baseline(zav1, wf, bsm, synth, imgs, S, qidx, qp, lam, args):
    """The same encode on the CPU over a bounded sample: 16 crops of 512x512
    (the top-left 2048x2048 of the first image), each decided as its own
    frame by the oracle's sequential restatement (or_wf_frame_levels) and
    written by the host still writer (one tile, one thread), one crop per
    thread, the threads pinned to the first CPUs this process may use; one
    untimed warm-up, then the median of 5 timed runs."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    C = 512
    crops = [np.ascontiguousarray(imgs[0, y:y + C, x:x + C]) for y in range(0, 4 * C, C) for x in range(0, 4 * C, C)]
    cfg = wf.funnel_cfg(C, C, 10, qidx, lam, 4, qp, n_images=1, leaf=16)
    model, cpus = cpu_info()
    threads = min(16, len(cpus), len(crops))
    keep = os.sched_getaffinity(0)
    os.sched_setaffinity(0, set(cpus[:threads]))
    try:
        def one(c):
            _, dec, lev = O.wf_frame(cfg, c[None], None, levels=True)
            leaves, levels = bsm.leaves_uniform(dec[0], lev[0], 16, cfg[0]['cands'])
            rc, stream = bsm.write_still(zav1, C, C, 10, 64, (1, 1), qidx, leaves, levels, threads=1)
            if rc:
                raise RuntimeError('zav1_write_still: %d' % rc)
            return stream
        times = []
        with ThreadPoolExecutor(threads) as ex:
            for rep in range(6):
                t0 = time.perf_counter()
                list(ex.map(one, crops))
                if rep:
                    times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, keep)
    med = float(np.median(times))
    return {'value': round(len(crops) * C * C / 1e6 / med, 4), 'unit': 'MP/s', 'cores': threads, 'kind': 'port',
            'sample': '%d crops of %dx%d (the top-left 2048x2048 of image 0), each decided as its own frame by '
                      'the oracle C restatement and written by the host still writer, one crop per thread; '
                      '1 warm-up + median of 5' % (len(crops), C, C),
            'seconds_median': round(med, 3), 'seconds_all': [round(t, 3) for t in times], 'cpu_model': model,
            'pinned_cpus': cpus[:threads], 'compiler': 'gcc ' + (oracle_cflags() or '?')}


def wavefront_headline(zav1, synth, args):
    wf = _load('zav1_wavefront', os.path.join(PKG, 'wavefront.py'))
    L = zav1.lib()
    S, n_img, qp = args.size, args.images, 40
    qidx = synth.qp_to_qindex(qp)
    lam = int(L.zav1_kf_full_lambda(qidx, qp, 10))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(min(16, len(os.sched_getaffinity(0)))) as ex:
        imgs = np.stack(list(ex.map(lambda i: synth.config2_luma(S, S, seed=2000 + i), range(n_img))))
    bsm = _load('zav1_bitstream', os.path.join(PKG, 'bitstream.py'))
    tiles = (bsm.legal_tile_cols(S, S, 64), 1)  # AV1's MAX_TILE_AREA: 2 tile columns at 4096 x 4096
    cfg = wf.funnel_cfg(S, S, 10, qidx, lam, 4, qp, n_images=n_img, leaf=16, tiles=tiles)  # preset 4: k1 15, k3 4
    md = wf.WavefrontMD(zav1, cfg)
    md.upload(imgs)
    st = zav1.stream_create()
    if args.profile_pass:  # one eager pass and nothing else, for profilers
        md.run(False, st)
        zav1.stream_sync(st)
        out = {'profile_pass': 'wavefront MD, %d x %dx%d 10-bit, eager' % (n_img, S, S), 'steps': md.info.n_steps,
               'leaves': md.info.n_leaves}
        md.close()
        return out
    use_graph = not args.no_graph
    for _ in range(max(1, args.warmup)):
        md.run(use_graph, st)
    zav1.stream_sync(st)
    # the plan's autotuner: the trellis form per leaf size, timed on a window
    # of its own steps (forms give identical levels); then the graph again
    tuned = md.autotune(stream=st)
    md.run(use_graph, st)
    zav1.stream_sync(st)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        md.run(use_graph, st)
    zav1.stream_sync(st)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000.0 / args.steps
    mp = n_img * S * S / 1e6
    value = mp / (ms_step / 1000.0)
    rec, dec = md.recon(), md.raw_decisions()
    # the eager form of the same chain must give the same bytes (not run
    # with --no-graph, which keeps graph capture out of profiled runs)
    same_forms = None
    if use_graph:
        md.run(False, st)
        zav1.stream_sync(st)
        same_forms = bool(np.array_equal(md.recon(), rec) and np.array_equal(md.raw_decisions(), dec))
    # per-stage device time of one eager pass (HIP events on the chain's stream),
    # and the trellis jobs it ran (the plan's counter, read around the pass)
    runs = np.zeros(2, np.uint64)
    zav1.check(L.zav1_memcpy_d2h(runs[0:].ctypes.data, md.info.d_trellis_runs, 8, st), 'runs')
    zav1.stream_sync(st)
    stage_ms = md.run_timed(st)
    zav1.check(L.zav1_memcpy_d2h(runs[1:].ctypes.data, md.info.d_trellis_runs, 8, st), 'runs')
    zav1.stream_sync(st)
    trellis_jobs = int(runs[1] - runs[0])
    assert np.array_equal(md.raw_decisions(), dec)
    info = md.info
    n_launch = info.n_steps  # one launch of each kernel per step (every step holds leaves)
    # plain per-call form on a bounded sample: the first P leaves in coding
    # order, recomputed from the decided neighbours; their records must come
    # back equal to the batched form's
    P = args.plain_leaves
    mark = dec.copy()
    mark.view(np.uint8)[:] = 0xFF
    md.set_raw_decisions(mark)
    tp0 = time.perf_counter()
    md.run_plain(0, P, st)
    zav1.stream_sync(st)
    tp1 = time.perf_counter()
    got = md.raw_decisions()
    redone = np.nonzero(got.view(np.uint8).reshape(len(got), -1).min(axis=1) != 0xFF)[0]
    plain_same = bool(len(redone) == P and np.array_equal(got[redone], dec[redone])
                      and np.array_equal(md.recon(), rec))
    md.set_raw_decisions(dec)
    plain_s_leaf = (tp1 - tp0) / P
    plain_mp_s = 16 * 16 / 1e6 / plain_s_leaf
    # roofline of the dominant stage's kernel
    dom = max(stage_ms, key=stage_ms.get)
    dom_s = stage_ms[dom] / 1000.0 / n_launch
    n_leaves = info.n_leaves
    n_types = info.n_types
    if dom == 'mds0_satd':
        ops = md_fast_ops_per_unit(wf.preset_candidates(), 16, 16) * n_leaves / n_launch
        roof = {'bound': 'valu', 'kernel': 'k_md_fast_satd<uint16_t,16,16>', 'unit': 'Tops/s',
                'achieved': round(ops / dom_s / 1e12, 4), 'peak': VALU_PEAK_TOPS,
                'frac': round(ops / dom_s / 1e12 / VALU_PEAK_TOPS, 5),
                'traffic': measured_traffic('k_md_fast_satd<unsigned short, 16, 16>'),
                'algorithmic_ops_per_launch': int(ops)}
    else:
        # the trellis: per block it walks, the block's transform coefficients,
        # levels and dequantized values are read (12 B per coefficient); per
        # job handed to it, the job record and eob (26 B).  Jobs = leaves x
        # survivors x tx types; the blocks walked are those the coefficient-
        # SATD gate passes with a nonzero eob (the plan's counter)
        jobs = n_leaves * 4 * n_types
        per_launch = (trellis_jobs * 256 * 12 + jobs * 26) / n_launch
        gbs = per_launch / dom_s / 1e9
        # the trellis kernel the autotuner kept for the 16x16 leaves
        form = tuned['forms'].get(16, 'queued')
        kname, kprefix = {'queued': ('k_rdoq_queue<TX_16X16, 16, 3>', 'k_rdoq_queue<2, 16, 3>'),
                          'direct': ('k_rdoq_group<TX_16X16, 16, 3, true>', 'k_rdoq_group<2, 16, 3, true>'),
                          'lane': ('k_rdoq<TX_16X16>', 'k_rdoq<2>')}[form]
        roof = {'bound': 'hbm', 'kernel': kname if dom == 'mds3_rdoq' else dom,
                'unit': 'GB/s', 'achieved': round(gbs, 3), 'peak': HBM_PEAK_GBS, 'frac': round(gbs / HBM_PEAK_GBS, 6),
                'traffic': measured_traffic(kprefix), 'algorithmic_bytes_per_launch': int(per_launch),
                'blocks_walked_per_launch': round(trellis_jobs / n_launch, 1),
                'jobs_per_launch': round(jobs / n_launch, 1),
                'note': 'latency-bound: a sequential trellis walk per block (the eob phase), 16 lanes per block; '
                        'the %s form, kept by the autotuner' % form}
    roof['stage'] = dom
    roof['avg_launch_us'] = round(dom_s * 1e6, 2)
    roof['launches'] = n_launch
    out = {
        'metric': 'MP/s encoded (AV1 still streams of the luma plane: superblock-wavefront intra mode decision on '
                  'the device, range coder on host threads, pipelined), batched vs plain per-call GPU',
        'value': None, 'unit': 'MP/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32', 'data': 'synthetic',
        'config': {'workload': 'encode: %d x %dx%d 10-bit luma in flight, qp 40, preset-4 funnel '
                               '(62 candidates, 15 -> 4 survivors, 6 tx-type groups, RDOQ), 16x16 leaves, SB64, '
                               '%dx%d tiles (the fewest within AV1\'s tile-area limit); mode decision as %s, '
                               'monochrome AV1 still streams written on host threads while the device decides the '
                               'next batch' % (n_img, S, S, tiles[0], tiles[1],
                                               'one HIP graph per batch' if use_graph else 'eager'),
                   'images_in_flight': n_img, 'width': S, 'height': S, 'bit_depth': 10, 'qp': qp, 'qindex': qidx,
                   'lambda': lam, 'leaf': 16, 'candidates': 62, 'k1': 15, 'k3': 4, 'tx_types': n_types,
                   'tile_cols': tiles[0], 'tile_rows': tiles[1],
                   'wavefront_steps': info.n_steps, 'max_leaves_per_step': info.max_leaves_per_step,
                   'device_bytes': int(info.bytes)},
        'md_only': {'value': round(value, 2), 'unit': 'MP/s', 'ms_per_step': round(ms_step, 3),
                    'what': 'the mode decision alone (graph replays of one plan, no streams)'},
        'forms_identical': {'graph_vs_eager': same_forms, 'plain_sample_vs_batched': plain_same},
        'autotune': tuned,
        'stage_ms': {k: round(v, 3) for k, v in stage_ms.items()},
        'eager_timed_pass_ms': round(sum(stage_ms.values()), 3),
        'roofline': roof,
        'plain_gpu': {'value': round(plain_mp_s, 4), 'unit': 'MP/s',
                      'sample': 'the first %d leaves in coding order (one launch per elementary call, selections '
                                'on the host)' % P,
                      'ms_per_leaf': round(plain_s_leaf * 1e3, 3),
                      'batched_over_plain': round(value / plain_mp_s, 1)},
    }
    if not args.no_cpu:
        out['cpu_baseline'] = wavefront_cpu_baseline(zav1, wf, bsm, synth, imgs, S, qidx, qp, lam, args)
    md.close()
    zav1.stream_destroy(st)
    if args.no_encode:  # decision only: the headline is the mode decision's rate
        out['metric'] = 'MP/s through the superblock-wavefront intra mode decision (luma), batched vs plain ' \
                        'per-call GPU (no streams: --no-encode)'
        out['value'], out['ms_per_step'] = round(value, 2), round(ms_step, 3)
        return out
    enc = encoded_stage(zav1, wf, imgs, cfg, args)
    out['value'], out['ms_per_step'] = enc.pop('value'), enc.pop('ms_per_step')
    out['encode'] = enc
    return out


def encoded_stage(zav1, wf, imgs, cfg, args):
    """The headline batch encoded: the same wavefront mode decision with the
    winners' levels kept, then the AV1 still stream of every image written
    on the host (zav1_wf_write_stills: the serial range coder per tile,
    whole images on parallel host threads).  Two plans alternate so the
    host writes batch k while the device decides batch k + 1 (the next
    batch's graph waits on an event for the previous one, so the device
    runs one batch at a time); a step is one batch decided and written.
    The two plans' streams must be byte-identical, and every timed batch's
    to the first."""
    c = np.array(cfg, copy=True)
    c[0]['keep_levels'] = 1
    L = zav1.lib()
    n_img, S = int(c[0]['n_images']), int(c[0]['width'])
    threads = min(16, len(os.sched_getaffinity(0)))
    plans = [wf.WavefrontMD(zav1, c) for _ in range(2)]
    sts = [zav1.stream_create() for _ in range(2)]
    evs = [zav1.event_create() for _ in range(2)]
    for p, st in zip(plans, sts):
        p.upload(imgs, st)
        zav1.stream_sync(st)
    ref = []
    for p, st in zip(plans, sts):
        p.run(True, st)
        p.autotune(stream=st)  # the trellis form per leaf size (identical results)
        for _ in range(max(1, args.warmup)):
            p.run(True, st)
            got = p.write_stills(threads=threads, stream=st)
        ref.append(got)
    nbytes = sum(len(x) for x in ref[0])
    plans_identical = ref[0] == ref[1]
    cap = 2 * nbytes + (1 << 20)
    # host-only cost of one batch's read-back and writing (device idle)
    tw0 = time.perf_counter()
    plans[0].write_stills(threads=threads, stream=sts[0], cap=cap)
    write_ms = (time.perf_counter() - tw0) * 1000.0
    K = args.steps
    repeat_identical = True
    t0 = time.perf_counter()
    plans[0].run(True, sts[0])
    zav1.event_record(evs[0], sts[0])
    for k in range(K):
        p, q = k % 2, (k + 1) % 2
        if k + 1 < K:
            zav1.check(L.zav1_stream_wait_event(sts[q], evs[p]), 'stream_wait_event')
            plans[q].run(True, sts[q])
            zav1.event_record(evs[q], sts[q])
        got = plans[p].write_stills(threads=threads, stream=sts[p], cap=cap)
        repeat_identical &= got == ref[0]
    t1 = time.perf_counter()
    for p in plans:
        p.close()
    for st in sts:
        zav1.stream_destroy(st)
    for e in evs:
        zav1.event_destroy(e)
    ms = (t1 - t0) * 1000.0 / K
    mp = n_img * S * S / 1e6
    return {'value': round(mp / (ms / 1000.0), 2), 'ms_per_step': round(ms, 3), 'steps': K,
            'host_threads': threads, 'host_write_ms_per_batch': round(write_ms, 2),
            'stream_bytes_per_batch': nbytes, 'bits_per_pixel': round(8.0 * nbytes / (n_img * S * S), 4),
            'streams_identical': {'two_plans': plans_identical, 'every_timed_batch': bool(repeat_identical)}}


def txsearch_stage(zav1, synth, tsm, opsd, args):
    """Exhaustive DC-predicted tx-type search over the 2048x2048 10-bit
    config-2 frame at qp 20 and 40 (19 sizes x the intra tx-type sets, one
    fused launch per size and quantizer pair)."""
    L = zav1.lib()
    S = 2048
    plane = synth.config2_luma(S, S, seed=2000)
    qps = (20, 40)
    qis = [synth.qp_to_qindex(q) for q in qps]
    sweep = tsm.TxSearchSweep(zav1, plane, 10, qis)
    n_cand = sweep.candidates_per_step()
    stream = zav1.stream_create()
    for _ in range(3):
        sweep.run(stream)
    zav1.sync(stream)
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        L.zav1_event_create(ctypes.byref(e))
    per_size_ms = {}
    steps = 10
    for ts in range(19):
        L.zav1_event_record(ev[0].value, stream)
        for _ in range(steps):
            sweep.run(stream, sizes=(ts,))
        L.zav1_event_record(ev[1].value, stream)
        zav1.sync(stream)
        ms = ctypes.c_float()
        L.zav1_event_elapsed_ms(ev[0].value, ev[1].value, ctypes.byref(ms))
        per_size_ms[ts] = ms.value / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        sweep.run(stream)
    zav1.sync(stream)
    ms_step = (time.perf_counter() - t0) * 1000.0 / steps
    dom = max(per_size_ms, key=per_size_ms.get)
    types = tsm.legal_types(dom)
    w, h = tsm.TX_DIMS[dom]
    nblk = (S // w) * (S // h)
    n_launch = tsm.launches(dom, len(qps))
    q_per_launch = len(qps) // n_launch
    avg_launch_s = per_size_ms[dom] / 1000.0 / n_launch
    ops_launch = nblk * block_ops(tsm, dom, types, opsd, q_per_launch)
    kname = {'lane': 'k_txsearch_lane', 'rows': 'k_txsearch_rows', 'group': 'k_txsearch'}[tsm.kernel_form(dom)]
    return {'workload': 'exhaustive DC-predicted tx-type search: 2048x2048 10-bit luma (config 2), qp 20 + 40, 19 '
                        'tx sizes x the intra ext-tx sets, fused DC-pred/fwd/fold/fp-quant/dist',
            'value': round(S * S / 1e6 / (ms_step / 1000.0), 2), 'unit': 'MP/s (both quantizers per pixel)',
            'ms_per_step': round(ms_step, 3), 'candidates_per_step': n_cand,
            'per_size_ms': {str(tsm.TX_DIMS[t]): round(per_size_ms[t], 4) for t in range(19)},
            'roofline': {'bound': 'valu', 'kernel': '%s<%d,%d>' % (kname, w, h), 'unit': 'Tops/s',
                         'achieved': round(ops_launch / avg_launch_s / 1e12, 3), 'peak': VALU_PEAK_TOPS,
                         'frac': round(ops_launch / avg_launch_s / 1e12 / VALU_PEAK_TOPS, 4),
                         'avg_launch_us': round(avg_launch_s * 1e6, 2),
                         'algorithmic_ops_per_launch': int(ops_launch)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--size', type=int, default=4096)
    ap.add_argument('--images', type=int, default=64)
    ap.add_argument('--no-graph', action='store_true', help='eager launches instead of the graph (profiling)')
    ap.add_argument('--profile-pass', action='store_true',
                    help='one eager pass of the headline workload only (for rocprofv3 runs)')
    ap.add_argument('--plain-leaves', type=int, default=48)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-encode', action='store_true', help='skip the encoded (bitstream) leg')
    ap.add_argument('--extras', default='lf,rdoq,serving,config4,preset8',
                    help='comma list of lf, rdoq, txsearch, config4, serving, preset8')
    ap.add_argument('--p8-images', type=int, default=16, help='images per batch of the preset-8 leg')
    ap.add_argument('--p8-profile-pass', action='store_true',
                    help='one batch of the preset-8 leg only (for rocprofv3 runs)')
    ap.add_argument('--copy-calibration', action='store_true')
    ap.add_argument('--pmc-summary', nargs=4, metavar=('FETCH_DIR', 'WRITE_DIR', 'CAL_FETCH_DIR', 'CAL_WRITE_DIR'),
                    help='print the per-kernel HBM traffic summary of rocprofv3 --pmc CSV runs (merged over '
                         'profiles/pmc_traffic.json)')
    ap.add_argument('--lf-steps', type=int, default=10)
    ap.add_argument('--streams', type=int, default=2)
    ap.add_argument('--serve-n', type=int, default=48)
    ap.add_argument('--serve-active', type=int, default=8)
    ap.add_argument('--serve-chunk', type=int, default=32)
    ap.add_argument('--serve-budget-gb', type=int, default=8)
    ap.add_argument('--serve-load', type=float, default=0.6)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU; multi-GPU runs are not part of it '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.size % 64 or args.size < 512 or args.images < 1 or args.steps < 1 or args.plain_leaves < 1:
        print('bench.py: --size must be a multiple of 64 from 512, --images/--steps/--plain-leaves >= 1',
              file=sys.stderr)
        sys.exit(2)

    if args.pmc_summary:
        prev = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
        print(json.dumps(pmc_summary(*args.pmc_summary, merge_into=prev if os.path.exists(prev) else None),
                         indent=1))
        return
    zav1 = _load('zav1', os.path.join(PKG, '__init__.py'))
    if args.copy_calibration:
        zav1.init(0)
        copy_calibration(zav1)
        return
    synth = _load('zav1_synth', os.path.join(PKG, 'synth.py'))
    zav1.init(0)
    if args.p8_profile_pass:
        print(json.dumps(preset8_stage(zav1, synth, args)))
        return
    out = wavefront_headline(zav1, synth, args)
    if args.profile_pass:
        print(json.dumps(out))
        return
    extras = [e for e in args.extras.split(',') if e]
    cpu = not args.no_cpu
    if 'txsearch' in extras:
        tsm = _load('zav1_txsearch', os.path.join(PKG, 'txsearch.py'))
    if 'lf' in extras:
        out['loopfilter_stage'] = loopfilter_stage(zav1, synth, args, cpu=cpu)
    if 'rdoq' in extras:
        out['rdoq'] = rdoq_stage(zav1, synth, args, cpu=cpu)
    if 'txsearch' in extras:
        opsd = json.load(open(os.path.join(PKG, 'csrc', 'txfm1d_ops.json')))
        out['txsearch_stage'] = txsearch_stage(zav1, synth, tsm, opsd, args)
    if 'config4' in extras:
        out['config4'] = config4_stage(zav1, synth, args)
    if 'serving' in extras:
        out['serving'] = serving_stage(zav1, synth, args)
    if 'preset8' in extras:
        out['preset8'] = preset8_stage(zav1, synth, args)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
