"""Benchmark: batched candidate RD (tx-type search) on one MI355X.

Workload (one step): the 2048x2048 10-bit 4:2:0 synthetic still (luma: 1/f^2
spectral noise in [64, 940] + 30 hard lines, seed 2000), at qp 20 and qp 40
(qindex 80 / 160).  For each qp, all 19 transform sizes tile the frame and
every block is evaluated with every transform type legal for its size
(DC prediction from the source neighbours -> residual -> forward transform
-> 64-point fold -> fp quantize -> distortion / eob / level sum), one fused
launch per (size, qp): 38 launches per step, ~4.4 M candidates.

value = luma megapixels x qps swept per second (inputs resident in HBM).
Also reported: the plain per-call GPU form (host pointers, one launch per
function-table call: fwd + quantize + sse per candidate) on a bounded sample,
the roofline of the dominant kernel from HIP events over the timed region, and
the oracle (CPU restatement) timed on a bounded crop of the same workload.

A third object, md_funnel, times the intra mode-decision funnel's MDS0 ->
MDS1 chain over every 16x16 unit of the config2 frame (61 candidates per
unit: prediction, Hadamard SATD, fast cost and exchange-sort top-3 on the
device, DCT_DCT quantize_b transform units of the survivors).

A second object, loopfilter_stage, times the loop-filter stage on the
4096x4096 8-bit 4:2:0 reconstruction (1/f noise + 30% 8x8 mean-flattened
blocks, seed 3000): deblocking of all planes, CDEF direction + 64-strength SSE
tables per 64x64 block, Wiener M/H + unit SSE and self-guided statistics per
restoration unit (128 luma / 64 chroma), with its own roofline and oracle
baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def md_funnel_stage(zav1, synth, args, cpu=True):
    """MDS0 -> MDS1 of the intra funnel over every 16x16 unit of the 2048x2048
    10-bit frame: 61 candidates per unit (13 modes + angle deltas), SATD,
    fast cost + exchange-sort survivors (k = 3), DCT_DCT quantize_b units."""
    L = zav1.lib()
    fm = _load('zav1_funnel', os.path.join(PKG, 'funnel.py'))
    S, bw = 2048, 16
    src = synth.config2_luma(S, S, seed=2000)
    rng = np.random.default_rng(2001)
    recon = np.clip(src.astype(np.int64) + rng.integers(-4, 5, src.shape), 0, 1023).astype(np.uint16)
    cands = fm.angle_candidates()
    nc = len(cands)
    rates = rng.integers(100, 1500, (1, nc)).astype(np.uint32)
    qi = synth.qp_to_qindex(40)
    lam, k = 90000, 3
    f = fm.IntraFunnel(zav1, src, recon, 10, bw, bw, qi, k=k, cands=cands, lam=lam, rates=rates)
    sp = ctypes.c_void_p()
    zav1.check(L.zav1_stream_create(ctypes.byref(sp)), 'stream_create')
    stream = sp.value
    for _ in range(2):
        f.run(stream)
    zav1.sync(stream)
    phases = f.stages()
    ev = []
    for _ in range(len(phases) + 1):
        e = ctypes.c_void_p()
        L.zav1_event_create(ctypes.byref(e))
        ev.append(e.value)
    phase_ms = np.zeros(len(phases))
    steps = args.lf_steps
    t0 = time.perf_counter()
    for _ in range(steps):
        L.zav1_event_record(ev[0], stream)
        for i, ph in enumerate(phases):
            f.run(stream, stages=(ph,))
            L.zav1_event_record(ev[i + 1], stream)
        zav1.sync(stream)
        for i in range(len(phases)):
            ms = ctypes.c_float()
            L.zav1_event_elapsed_ms(ev[i], ev[i + 1], ctypes.byref(ms))
            phase_ms[i] += ms.value
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000.0 / steps
    phase_ms /= steps
    n = bw * bw
    npred = f.predictions()
    ntx = f.n_tx
    nu = f.n_units
    # fused fast loop: algorithmic int ops per launch (per candidate sample:
    # prediction by mode class + residual + Hadamard/SATD), and compulsory
    # HBM bytes (source block + reference samples per unit, u32 SATD per candidate)
    ops_unit = md_fast_ops_per_unit(cands, bw, bw)
    satd_bytes = nu * (2 * n + (2 * bw + 2 * bw + 1) * 2 + 20) + npred * 4
    dom = phases[int(np.argmax(phase_ms))]
    dom_s = phase_ms[phases.index(dom)] / 1000.0
    if dom == 'satd':
        roof = {'bound': 'valu', 'kernel': 'k_md_fast_satd<uint16_t,16>', 'unit': 'Tops/s',
                'achieved': round(nu * ops_unit / dom_s / 1e12, 3), 'peak': VALU_PEAK_TOPS,
                'frac': round(nu * ops_unit / dom_s / 1e12 / VALU_PEAK_TOPS, 4),
                'traffic': measured_traffic('k_md_fast_satd<unsigned short, 16, 16>'),
                'avg_launch_us': round(dom_s * 1e6, 2), 'algorithmic_ops_per_launch': int(nu * ops_unit),
                'hbm_algorithmic_GBs': round(satd_bytes / dom_s / 1e9, 2)}
    else:
        bytes_of = {'pred': nu * k * (2 * n + (4 * bw + 1) * 2 + 20),
                    'select': nu * nc * 8 + ntx * 16 + nu * k * 9,
                    'tx': ntx * (2 * n + 2 * n + 16 + 13)}
        gbs = bytes_of[dom] / dom_s / 1e9
        roof = {'bound': 'hbm', 'kernel': {'pred': 'k_intra_pred<uint16_t,SurvivorJobs>', 'select': 'k_fast_select',
                                           'tx': 'k_txunit<16,16,uint16_t>'}[dom],
                'unit': 'GB/s', 'achieved': round(gbs, 1), 'peak': HBM_PEAK_GBS, 'frac': round(gbs / HBM_PEAK_GBS, 4),
                'traffic': None, 'avg_launch_us': round(dom_s * 1e6, 2),
                'algorithmic_bytes_per_launch': int(bytes_of[dom])}
    out = {'workload': 'intra funnel MDS0 -> MDS1: 2048x2048 10-bit luma (config2), all 16x16 units, 61 '
                       'candidates each (13 modes + angle deltas, synthetic neighbour canvas and mode rates), '
                       'fused prediction + Hadamard SATD on chip, fast cost + exchange-sort top-3, survivors '
                       'predicted again, DCT_DCT quantize_b units at qp 40',
           'value': round(S * S / 1e6 / (ms_step / 1000.0), 2), 'unit': 'MP/s', 'steps': steps,
           'ms_per_step': round(ms_step, 3), 'candidates_per_step': int(npred), 'tx_units_per_step': int(ntx),
           'phase_ms': {ph: round(float(phase_ms[i]), 4) for i, ph in enumerate(phases)},
           'roofline': roof}
    # the two-pass form (every prediction through HBM), same results, for comparison
    f2 = fm.IntraFunnel(zav1, src, recon, 10, bw, bw, qi, k=k, cands=cands, lam=lam, rates=rates, fused=False)
    f2.run(stream)
    zav1.sync(stream)
    t2 = time.perf_counter()
    for _ in range(steps):
        f2.run(stream)
    zav1.sync(stream)
    out['two_pass_ms_per_step'] = round((time.perf_counter() - t2) * 1000.0 / steps, 3)
    del f2
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        c = S
        xy = fm.unit_grid(c, c, bw, bw)
        jobs = np.zeros(len(xy) * nc, O.INTRA_JOB)
        for u, (x, y) in enumerate(xy.tolist()):
            nt, ntr, nl, nbl = fm.raster_availability(x, y, bw, bw, c)
            for m, (mode, delta, fi) in enumerate(cands):
                jobs[u * nc + m] = (x, y, bw, bw, mode, delta, fi, 1, 0, 0, nt, ntr, nl, nbl, 0, 0)
        threads = min(os.cpu_count() or 1, 16)
        os.environ['OMP_NUM_THREADS'] = str(threads)
        s_c, r_c = np.ascontiguousarray(src[:c, :c]), np.ascontiguousarray(recon[:c, :c])
        tc0 = time.perf_counter()
        O.md_funnel(s_c, r_c, 10, jobs, nc, rates, lam, k, f.tx_size, 0, qi)
        tc1 = time.perf_counter()
        out['cpu_baseline'] = {'value': round(c * c / 1e6 / (tc1 - tc0), 4), 'unit': 'MP/s', 'cores': threads,
                               'kind': 'port', 'sample': '%dx%d (whole) frame, same chain (oracle C, OpenMP '
                                                         'over units)' % (c, c),
                               'seconds': round(tc1 - tc0, 2)}
    return out


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

    def launch():
        zav1.rdoq_batch(ts, qi, 10, d_coef, d_q, d_dq, d_eob, d_jobs, nj, d_costs, d_scans, stream=stream)

    steps = max(5, args.steps // 5)
    for _ in range(2):
        restore()
        launch()
    zav1.sync(stream)
    tot_ms = 0.0
    for _ in range(steps):
        restore()
        L.zav1_event_record(ev[0].value, stream)
        launch()
        L.zav1_event_record(ev[1].value, stream)
        zav1.sync(stream)
        ms = ctypes.c_float()
        L.zav1_event_elapsed_ms(ev[0].value, ev[1].value, ctypes.byref(ms))
        tot_ms += ms.value
    ms_step = tot_ms / steps
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
           'roofline': {'bound': 'latency', 'kernel': 'k_rdoq<TX_16X16>', 'unit': 'GB/s',
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


def config4_stage(zav1, synth, tsm, args, streams):
    """Config 4 at its per-GPU shape: one tile column (2048 wide, 4096 high)
    of the 8192x4096 10-bit frame with screen regions, tx-type search at
    qp 28 (one quantizer), sizes spread over the bench's streams."""
    plane = synth.config4_tile_column(4096, 2048, seed=4000)
    sw = tsm.TxSearchSweep(zav1, plane, 10, [synth.qp_to_qindex(28)])
    for _ in range(2):
        sw.run(streams)
    for sx in streams:
        zav1.sync(sx)
    steps = max(5, args.steps // 5)
    t0 = time.perf_counter()
    for _ in range(steps):
        sw.run(streams)
    for sx in streams:
        zav1.sync(sx)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    mp = plane.shape[0] * plane.shape[1] / 1e6
    return {'metric': 'MP/s (luma megapixels x qps) through the tx-type search', 'value': round(mp / ms * 1000.0, 2),
            'unit': 'MP/s', 'ms_per_step': round(ms, 3), 'steps': steps,
            'config': {'workload': 'config4 tile column: 2048x4096 10-bit luma (1/f noise, 25% screen regions), '
                                   'qp 28, 19 tx sizes x legal types', 'streams': len(streams)}}


def serving_stage(zav1, synth, tsm, args):
    """Serving mix at a fixed Poisson arrival rate: per-request latency of
    the tx-type search stage (upload, grids, every size at the request's
    qp, results resident in HBM) with an HBM admission pool of contexts.
    The rate is set to args.serve_load of the pool's measured back-to-back
    throughput."""
    srv = _load('zav1_serving', os.path.join(PKG, 'serving.py'))
    reqs = synth.serving_mix(args.serve_n, seed=5000)
    imgs = synth.serving_images(reqs, seed=5000)
    loop = srv.ServingLoop(zav1, tsm, synth, n_ctx=args.serve_ctx)
    # back-to-back throughput (every request arrives at once), twice: the
    # first pass warms up
    for _ in range(2):
        lat0, mp0, wall0 = loop.run(reqs, imgs, rate_hz=1e9)
    mp_per_s = mp0 / wall0
    rate = args.serve_load * mp_per_s / (mp0 / len(reqs))
    lat, mp, wall = loop.run(reqs, imgs, rate_hz=rate, seed=5002)
    loop.close()
    return {'metric': 'per-image latency (ms) of the tx-type search stage, Poisson arrivals',
            'p50_ms': round(float(np.percentile(lat, 50)) * 1e3, 3),
            'p99_ms': round(float(np.percentile(lat, 99)) * 1e3, 3),
            'mean_ms': round(float(lat.mean()) * 1e3, 3),
            'value': round(mp / wall, 2), 'unit': 'MP/s', 'requests': len(reqs), 'megapixels': round(mp, 2),
            'arrival_rate_hz': round(rate, 1), 'load': args.serve_load,
            'back_to_back_MPs': round(mp_per_s, 2),
            'admission': {'contexts': args.serve_ctx, 'hbm_per_context_MB': round(loop.budget / 1e6, 1)},
            'config': {'workload': 'serving mix: area log-uniform 256^2..4096x3072, aspect 1:1..16:9, 50/50 '
                                   '8/10-bit, qp U{20..50}; per request upload + tx-type search over every size '
                                   '(one quantizer)'}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--size', type=int, default=2048)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-lf', action='store_true')
    ap.add_argument('--no-funnel', action='store_true')
    ap.add_argument('--copy-calibration', action='store_true')
    ap.add_argument('--lf-steps', type=int, default=10)
    ap.add_argument('--streams', type=int, default=2)
    ap.add_argument('--no-config4', action='store_true')
    ap.add_argument('--no-serving', action='store_true')
    ap.add_argument('--no-rdoq', action='store_true')
    ap.add_argument('--serve-n', type=int, default=96)
    ap.add_argument('--serve-ctx', type=int, default=4)
    ap.add_argument('--serve-load', type=float, default=0.6)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU; multi-GPU runs are not part of it '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    zav1 = _load('zav1', os.path.join(PKG, '__init__.py'))
    if args.copy_calibration:
        zav1.init(0)
        copy_calibration(zav1)
        return
    synth = _load('zav1_synth', os.path.join(PKG, 'synth.py'))
    tsm = _load('zav1_txsearch', os.path.join(PKG, 'txsearch.py'))
    opsd = json.load(open(os.path.join(PKG, 'csrc', 'txfm1d_ops.json')))
    zav1.init(0)
    L = zav1.lib()

    S = args.size
    plane = synth.config2_luma(S, S, seed=2000)
    qps = (20, 40)
    qis = [synth.qp_to_qindex(q) for q in qps]
    sweep = tsm.TxSearchSweep(zav1, plane, 10, qis)
    n_cand = sweep.candidates_per_step()

    # explicit stream: kernels and events share it
    st = ctypes.c_void_p()
    zav1.check(L.zav1_stream_create(ctypes.byref(st)), 'stream_create')
    stream = st.value

    for _ in range(args.warmup):
        sweep.run(stream)
    zav1.sync(stream)

    def event():
        e = ctypes.c_void_p()
        zav1.check(L.zav1_event_create(ctypes.byref(e)), 'event_create')
        return e.value

    def elapsed(a, b):
        ms = ctypes.c_float()
        zav1.check(L.zav1_event_elapsed_ms(a, b, ctypes.byref(ms)), 'event_elapsed')
        return ms.value

    # per-size breakdown (untimed pass, HIP events around each size's launch)
    cal_steps = 3
    evs = {ts: (event(), event()) for ts in range(19)}
    per_size_ms = {ts: 0.0 for ts in range(19)}
    for _ in range(cal_steps):
        for ts in range(19):
            L.zav1_event_record(evs[ts][0], stream)
            sweep.run(stream, sizes=(ts,))
            L.zav1_event_record(evs[ts][1], stream)
        zav1.sync(stream)
        for ts in range(19):
            per_size_ms[ts] += elapsed(*evs[ts]) / cal_steps
    dom = max(per_size_ms, key=per_size_ms.get)

    # timed region: the steps back to back, the other sizes spread over
    # args.streams streams (kernels of different sizes overlap, filling each
    # other's tails).  The dominant size runs alone at the start of every
    # step (all streams join before it and fork after it), bracketed by HIP
    # events on its stream, so its measured duration is its own (read after
    # the region).
    streams = [stream]
    for _ in range(args.streams - 1):
        sx = ctypes.c_void_p()
        zav1.check(L.zav1_stream_create(ctypes.byref(sx)), 'stream_create')
        streams.append(sx.value)
    dom_ev = [(event(), event()) for _ in range(args.steps)]
    join_ev = [[event() for _ in streams] for _ in range(args.steps)]
    for sx in streams:
        zav1.sync(sx)
    t0 = time.perf_counter()
    for k in range(args.steps):
        for j, sx in enumerate(streams[1:], 1):  # join: the stream of the dominant waits for the others
            L.zav1_event_record(join_ev[k][j], sx)
            L.zav1_stream_wait_event(stream, join_ev[k][j])
        L.zav1_event_record(dom_ev[k][0], stream)
        sweep.run(stream, sizes=(dom,))
        L.zav1_event_record(dom_ev[k][1], stream)
        for sx in streams[1:]:  # fork: the others start after it
            L.zav1_stream_wait_event(sx, dom_ev[k][1])
        for ts in range(19):
            if ts != dom:
                sweep.run(streams, sizes=(ts,))
    for sx in streams:
        zav1.sync(sx)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    mp = S * S / 1e6 * len(qps)
    value = mp / (ms_per_step / 1000.0)
    dom_ms_total = sum(elapsed(*e) for e in dom_ev)

    # dominant kernel: the tx size with the largest event time.  Every form
    # evaluates both quantizers in one launch (transforms shared).
    types = tsm.legal_types(dom)
    w, h = tsm.TX_DIMS[dom]
    nblk = (S // w) * (S // h)
    form = tsm.kernel_form(dom)
    n_launch = tsm.launches(dom, len(qps))
    q_per_launch = len(qps) // n_launch
    avg_launch_s = dom_ms_total / 1000.0 / (args.steps * n_launch)
    ops_launch = nblk * block_ops(tsm, dom, types, opsd, q_per_launch)
    ops_launch_per_candidate = nblk * block_ops(tsm, dom, types, opsd) * q_per_launch
    bytes_launch = nblk * (w * h * 2 + (w + h) * 2 + len(types) * 16 * q_per_launch)
    achieved_tops = ops_launch / avg_launch_s / 1e12
    kname = {'lane': 'k_txsearch_lane', 'rows': 'k_txsearch_rows', 'group': 'k_txsearch'}[form]

    dom_pmc = pmc_kernel('%s<%d, %d, ' % (kname, w, h), bd=10, nq=q_per_launch)
    out = {
        'metric': 'MP/s (luma megapixels x qps) through batched candidate RD, byte-identical to the oracle',
        'value': round(value, 2), 'unit': 'MP/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 3), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32', 'data': 'synthetic',
        'config': {'workload': 'config2 tx-type search: %dx%d 10-bit luma, qp 20+40, 19 tx sizes x legal '
                               'types, fused DC-pred/fwd/fold/fp-quant/dist' % (S, S),
                   'candidates_per_step': n_cand,
                   'launches_per_step': sum(tsm.launches(t, len(qps)) for t in range(19)),
                   'streams': args.streams},
        'per_size_ms_per_step': {str(tsm.TX_DIMS[t]): round(per_size_ms[t], 4) for t in range(19)},
        'roofline': {'bound': 'valu', 'kernel': '%s<%d,%d>' % (kname, w, h), 'achieved': round(achieved_tops, 3),
                     'peak': VALU_PEAK_TOPS, 'unit': 'Tops/s', 'frac': round(achieved_tops / VALU_PEAK_TOPS, 4),
                     'traffic': measured_traffic(dom_pmc),
                     'valu_issue_frac': (None if measured_valu_insts(dom_pmc) is None else
                                         round(measured_valu_insts(dom_pmc) * 64 / avg_launch_s / VALU_ISSUE_PEAK,
                                               4)),
                     'pmc_kernel': dom_pmc,
                     'quantizers_per_launch': q_per_launch,
                     'avg_launch_us': round(avg_launch_s * 1e6, 2),
                     'algorithmic_ops_per_launch': int(ops_launch),
                     'ops_per_launch_transform_per_quantizer': int(ops_launch_per_candidate),
                     'hbm_algorithmic_GBs': round(bytes_launch / avg_launch_s / 1e9, 2),
                     'hbm_peak_GBs': HBM_PEAK_GBS},
    }

    if not args.no_plain:
        # plain per-call form: function table with host pointers, one launch per call
        tab = zav1.dsp_table()
        rng = np.random.default_rng(1)
        n = 64
        blocks = rng.integers(-1023, 1024, (n, 16, 16)).astype(np.int32)
        row = zav1.quant_row(qis[0], 10)
        rounding = np.array(list(row.round_fp), np.int32)
        qfp = np.array(list(row.quant_fp), np.int32)
        deq = np.array(list(row.dequant), np.int32)
        P = ctypes.POINTER(ctypes.c_int32)
        co = np.zeros(256, np.int32)
        q = np.zeros(256, np.int32)
        dq = np.zeros(256, np.int32)
        res = ctypes.c_uint64()
        tp0 = time.perf_counter()
        calls = 0
        for j in range(n):
            for t in range(16):
                tab.fwd_txfm2d(blocks[j].ctypes.data_as(P), co.ctypes.data_as(P), 16, 2, t)
                tab.quantize_fp_raster(co.ctypes.data_as(P), q.ctypes.data_as(P), dq.ctypes.data_as(P), 256,
                                       rounding.ctypes.data_as(P), qfp.ctypes.data_as(P),
                                       deq.ctypes.data_as(P), 0)
                tab.sse_i32(co.ctypes.data_as(P), dq.ctypes.data_as(P), 256, ctypes.byref(res))
                calls += 3
        tp1 = time.perf_counter()
        per_cand = (tp1 - tp0) / (n * 16)
        plain_ms_step = per_cand * n_cand * 1000.0
        out['plain_gpu'] = {'value': round(mp / (plain_ms_step / 1000.0), 4), 'unit': 'MP/s',
                            'sample': '%d 16x16 blocks x 16 types, 3 table calls each (%d calls), '
                                      'extrapolated per candidate to the step' % (n, calls),
                            'us_per_candidate': round(per_cand * 1e6, 2),
                            'batched_over_plain': round(value / (mp / (plain_ms_step / 1000.0)), 1)}

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        crop = S
        sub = np.ascontiguousarray(plane[:crop, :crop])
        threads = min(os.cpu_count() or 1, 16)
        os.environ['OMP_NUM_THREADS'] = str(threads)
        tc0 = time.perf_counter()
        for qi in qis:
            for ts in range(19):
                xy = tsm.tile_xy(crop, crop, ts)
                O.txsearch_plane(sub, 10, ts, xy, np.array(tsm.legal_types(ts), np.uint8), qi)
        tc1 = time.perf_counter()
        cpu_mp = crop * crop / 1e6 * len(qps) / (tc1 - tc0)
        out['cpu_baseline'] = {'value': round(cpu_mp, 4), 'unit': 'MP/s', 'cores': threads, 'kind': 'port',
                               'sample': '%dx%d (whole) frame, qp 20+40, all sizes/types '
                                         '(oracle C, OpenMP over blocks)' % (crop, crop),
                               'seconds': round(tc1 - tc0, 2)}
    if not args.no_config4:
        out['config4_tile'] = config4_stage(zav1, synth, tsm, args, streams)
    if not args.no_serving:
        out['serving'] = serving_stage(zav1, synth, tsm, args)
    if not args.no_lf:
        out['loopfilter_stage'] = loopfilter_stage(zav1, synth, args, cpu=not args.no_cpu)
    if not args.no_funnel:
        out['md_funnel'] = md_funnel_stage(zav1, synth, args, cpu=not args.no_cpu)
    if not args.no_rdoq:
        out['rdoq'] = rdoq_stage(zav1, synth, args, cpu=not args.no_cpu)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
