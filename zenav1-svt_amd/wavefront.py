"""Host driver of the superblock-wavefront mode decision (zav1_wf_*): the
batched form of the reference's leaf funnel over every leaf of a batch of
still images (rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs:1007
evaluate_leaf, in the coding order of enc/pipeline.rs:7588-7598).

`WavefrontMD` plans a batch (n_images equal-size luma planes, uniform
leaves, tiles), uploads sources into the stacked device plane, runs the
decision (eager launches or a replayed HIP graph, or the plain per-call
form for a range of leaves) and returns the reconstruction and the per-leaf
decisions.
"""
import ctypes

import numpy as np

CAND = np.dtype([('mode', 'u1'), ('delta', 'i1'), ('fi_mode', 'u1'), ('pad', 'u1')])
CFG = np.dtype([('width', '<i4'), ('height', '<i4'), ('n_images', '<i4'), ('bd', '<i4'), ('leaf', '<i4'),
                ('sb', '<i4'), ('tile_cols', '<i4'), ('tile_rows', '<i4'), ('qindex', '<i4'), ('k1', '<i4'),
                ('k3', '<i4'), ('rdoq_level', '<i4'), ('txt_groups', '<i4'), ('txt_rate_th', '<i4'),
                ('txt_satd_th', '<i4'), ('coeff_rate_est_lvl', '<i4'), ('lambda', '<u8'), ('n_cands', '<i4'), ('pad', '<i4'),
                ('cands', CAND, 64), ('txt_groups_lt16', '<i4'), ('keep_levels', '<i4'), ('leaf_map', '<u8')])
DECISION = np.dtype([('cand', 'u1'), ('tx_type', 'u1'), ('mode', 'u1'), ('fi_mode', 'u1'), ('eob', '<i4'),
                     ('rate', '<i4'), ('mode_rate', '<i4'), ('dist', '<u8'), ('cost', '<u8')])
MD_UNIT = np.dtype([('x', '<i4'), ('y', '<i4'), ('n_top_px', '<i2'), ('n_topright_px', '<i2'),
                    ('n_left_px', '<i2'), ('n_bottomleft_px', '<i2'), ('edge_filter', 'u1'), ('filt_type', 'u1'),
                    ('pad0', 'u1'), ('pad1', 'u1')])


class WfInfo(ctypes.Structure):
    _fields_ = [('d_src', ctypes.c_void_p), ('d_recon', ctypes.c_void_p), ('d_decisions', ctypes.c_void_p),
                ('d_units', ctypes.c_void_p), ('n_leaves', ctypes.c_int32), ('n_steps', ctypes.c_int32),
                ('max_leaves_per_step', ctypes.c_int32), ('n_types', ctypes.c_int32),
                ('stacked_height', ctypes.c_int32), ('types', ctypes.c_uint8 * 16), ('bytes', ctypes.c_size_t),
                ('d_levels', ctypes.c_void_p), ('d_level_off', ctypes.c_void_p), ('n_levels', ctypes.c_size_t),
                ('d_trellis_runs', ctypes.c_void_p)]


def preset_candidates(fi=True):
    """The intra candidates of the mode decision at presets 1-4 (mode_end
    PAETH, angular level 1: every delta of the 8 directional modes) with
    FILTER_DC the only filter-intra mode (fi_max 0): 13 + 48 + 1 = 62, in
    the reference's injection order (angle_candidates)."""
    return angle_candidates(1, 12, (0,) if fi else ())


def qp_scaled_satd_th(th, cli_qp):
    """div_round(th * w, d) with svt_aom_get_qp_based_th_scaling_factors
    (rust/crates/svtav1-encoder/src/pd0.rs:185, leaf_funnel/mod.rs:7210)."""
    import math
    if cli_qp >= 46:
        w, d = int((1.05 - math.exp(-(max(cli_qp, 40) - 35.0) / 10.0)) * 10000.0), 10000
    else:
        w, d = max(cli_qp, 10), 63
    return (th * w + d // 2) // d


def _div_round(a, b):
    return (a + b // 2) // b


def _qp_factors(cli_qp):
    """svt_aom_get_qp_based_th_scaling_factors (pd0.rs:185)."""
    import math
    if cli_qp >= 46:
        return int((1.05 - math.exp(-(max(cli_qp, 40) - 35.0) / 10.0)) * 10000.0), 10000
    return max(cli_qp, 10), 63


def angle_candidates(angular_level, mode_end=12, fi_modes=(), use_angle=True):
    """inject_intra_candidates (rust/crates/svtav1-encoder/src/leaf_funnel/
    mod.rs:1940-1975), in its order: modes DC..mode_end, each directional
    mode with its deltas in place; at angular level >= 4 the D45..D67 modes
    are masked; at levels <= 2 (and use_angle: blocks of 8x8 and up) every
    directional mode takes the deltas -3..3 (level 2: only -3, 0, 3); then
    the filter-intra modes (DC_PRED blocks carrying them)."""
    c = []
    for m in range(mode_end + 1):
        if 3 <= m <= 8 and angular_level >= 4:
            continue
        if 1 <= m <= 8 and angular_level <= 2 and use_angle:
            c += [(m, d, 5) for d in range(-3, 4) if not (angular_level >= 2 and d in (-2, -1, 1, 2))]
        else:
            c.append((m, 0, 5))
    return c + [(0, 0, f) for f in fi_modes]


# FunnelCfg::for_preset (rust/crates/svtav1-encoder/src/leaf_funnel/
# rate_tables.rs:660): preset -> (mode_end, angular level, filter-intra
# modes, nic num, txt groups < 16 / >= 16, SATD gate, rate gate)
_PRESETS = {1: (12, 1, (0,), 12, 6, 6, 20, 250), 2: (12, 1, (0,), 12, 6, 6, 20, 250),
            3: (12, 1, (0,), 6, 6, 6, 20, 250), 4: (12, 1, (0,), 6, 6, 6, 15, 250),
            5: (12, 2, (0,), 6, 6, 6, 15, 250), 6: (9, 4, (0,), 6, 5, 4, 10, 100),
            7: (9, 4, (), 4, 3, 2, 10, 50), 8: (9, 4, (), 0, 3, 2, 10, 50)}


def preset_funnel(preset, cli_qp, leaf=16):
    """The wavefront MD's funnel settings at a preset for one leaf size, from
    FunnelCfg::for_preset: the candidate list (angle_candidates; filter
    intra only on leaves up to 32), the NIC counts of MDS1 / MDS3
    (nic_counts, rate_tables.rs:1054: bases 64 / 16 x num / 16, at least 2
    when num > 0 else 1, qp-scaled; capped at the candidates), the tx-type
    groups, the rate and SATD gates (the SATD one qp-scaled) and the RDOQ
    level (rdoq_level_allintra, quant.rs:1535: 1 up to preset 5, above it by
    the picture's coefficient level -- taken as Normal, level 3) and the luma
    rate estimate (coeff_rate_est_lvl: 2, the small-eob closed form, at
    presets 7 and 8, rate_tables.rs:964).  Presets
    1..8 (preset 0's 66 candidates exceed the configuration's 64)."""
    if preset not in _PRESETS:
        raise ValueError('preset %r: 1..8' % (preset,))
    mode_end, ang, fi_modes, nic, g_lt16, g_ge16, satd, rate = _PRESETS[preset]
    cands = angle_candidates(ang, mode_end, fi_modes if leaf <= 32 else (), use_angle=leaf >= 8)
    qw, qd = _qp_factors(cli_qp)

    def scale(base, num):
        lo = 2 if num else 1
        n = max(lo, _div_round(base * num, 16))
        return max(lo, _div_round(n * qw, qd))
    k1 = min(scale(64, nic), len(cands))
    k3 = min(scale(16, nic), k1)
    return dict(cands=cands, k1=k1, k3=k3, txt_groups=g_ge16 if leaf >= 16 else g_lt16, txt_rate_th=rate,
                txt_satd_th=qp_scaled_satd_th(satd, cli_qp), rdoq_level=1 if preset <= 5 else 3,
                coeff_rate_est_lvl=2 if preset >= 7 else 1)


def funnel_cfg(width, height, bd, qindex, lam, preset, cli_qp, **kw):
    """make_cfg with every funnel setting of the preset (preset_funnel) at
    the configuration's leaf size."""
    f = preset_funnel(preset, cli_qp, kw.get('leaf', 16))
    return make_cfg(width, height, bd, qindex, lam, **f, **kw)


def tree_cfg(width, height, bd, qindex, lam, preset, cli_qp, leaf_map, **kw):
    """funnel_cfg for the leaves of a partition map (the partition pre-pass's
    leaf maps, n_images x units x 256): the candidates with filter intra
    (64x64 leaves drop it), the tx-type groups of leaves from 16x16 and of
    the smaller ones.  The map must stay alive until the plan is created."""
    f = preset_funnel(preset, cli_qp, 16)
    f_small = preset_funnel(preset, cli_qp, 8)
    cfg = make_cfg(width, height, bd, qindex, lam, **f, **kw)
    cfg[0]['txt_groups_lt16'] = f_small['txt_groups']
    if not (isinstance(leaf_map, np.ndarray) and leaf_map.dtype == np.uint8 and leaf_map.flags['C_CONTIGUOUS']):
        raise TypeError('leaf_map: a C-contiguous uint8 array (its address goes into the configuration)')
    cfg[0]['leaf_map'] = leaf_map.ctypes.data
    return cfg


def make_cfg(width, height, bd, qindex, lam, *, n_images=1, leaf=16, sb=64, tiles=(1, 1), k1=15, k3=4,
             rdoq_level=1, txt_groups=6, txt_rate_th=250, txt_satd_th=0, coeff_rate_est_lvl=1, cands=None):
    cands = preset_candidates() if cands is None else cands
    cfg = np.zeros(1, CFG)
    c = cfg[0]
    c['width'], c['height'], c['n_images'], c['bd'], c['leaf'], c['sb'] = width, height, n_images, bd, leaf, sb
    c['tile_cols'], c['tile_rows'], c['qindex'], c['k1'], c['k3'] = tiles[0], tiles[1], qindex, k1, k3
    c['rdoq_level'], c['txt_groups'], c['txt_rate_th'], c['txt_satd_th'] = rdoq_level, txt_groups, txt_rate_th, \
        txt_satd_th
    c['coeff_rate_est_lvl'] = coeff_rate_est_lvl
    c['lambda'], c['n_cands'] = lam, len(cands)
    for i, (m, d, f) in enumerate(cands):
        c['cands'][i] = (m, d, f, 0)
    return cfg


class WavefrontMD:
    def __init__(self, zav1, cfg):
        self.z = zav1
        self.cfg = np.ascontiguousarray(cfg, CFG)
        L = zav1.lib()
        h = ctypes.c_void_p()
        rc = L.zav1_wf_create(self.cfg.ctypes.data, ctypes.byref(h))
        if rc:
            raise zav1.Zav1Error(rc, 'zav1_wf_create (%s)' % L.zav1_wf_last_error().decode())
        self.h = h
        self.info = WfInfo()
        zav1.check(L.zav1_wf_info(h, ctypes.byref(self.info)), 'zav1_wf_info')
        c = self.cfg[0]
        self.W, self.H, self.n, self.bd = int(c['width']), int(c['height']), int(c['n_images']), int(c['bd'])
        self.dt = np.uint8 if self.bd == 8 else np.uint16

    @staticmethod
    def footprint(zav1, cfg):
        b = ctypes.c_size_t()
        zav1.check(zav1.lib().zav1_wf_footprint(np.ascontiguousarray(cfg, CFG).ctypes.data, ctypes.byref(b)),
                   'zav1_wf_footprint')
        return b.value

    def upload(self, images, stream=None):
        """images: (n_images, height, width) samples."""
        a = np.ascontiguousarray(images, self.dt).reshape(self.n * self.H, self.W)
        self.z.check(self.z.lib().zav1_memcpy_h2d(self.info.d_src, a.ctypes.data, a.nbytes, stream), 'upload')

    def run(self, use_graph=False, stream=None):
        rc = self.z.lib().zav1_wf_run(self.h, int(use_graph), stream)
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_run (%s)' % self.z.lib().zav1_wf_last_error().decode())

    STAGES = ('rates', 'mds0_satd', 'mds0_select', 'predict', 'mds1_jobs', 'mds1_units', 'mds3_jobs',
              'mds3_fwd_quant', 'mds3_rdoq', 'mds3_rate_dist_recon', 'commit')

    def run_timed(self, stream=None):
        """Eager run with HIP events at every stage boundary: {stage: total ms}."""
        out = (ctypes.c_double * len(self.STAGES))()
        rc = self.z.lib().zav1_wf_run_timed(self.h, stream, out, len(self.STAGES))
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_run_timed (%s)' % self.z.lib().zav1_wf_last_error().decode())
        return dict(zip(self.STAGES, list(out)))

    FORMS = ('queued', 'direct', 'lane')

    def autotune(self, first=None, count=None, stream=None):
        """The plan's trellis form per leaf size, timed on steps [first,
        first + count) (default: up to 256 steps from the middle of the
        plan): {'forms': {leaf: form name}, 'ms': {leaf: {form: ms}}}.  The
        decisions are unchanged (every form gives the same levels)."""
        n = self.info.n_steps
        count = min(256, n) if count is None else count
        first = max(0, (n - count) // 2) if first is None else first
        forms = (ctypes.c_int32 * 4)()
        ms = (ctypes.c_double * 12)()
        rc = self.z.lib().zav1_wf_autotune(self.h, first, count, stream, forms, ms)
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_autotune (%s)' % self.z.lib().zav1_wf_last_error().decode())
        out = {'forms': {}, 'ms': {}}
        for k, leaf in enumerate((8, 16, 32, 64)):
            if forms[k] >= 0:
                out['forms'][leaf] = self.FORMS[forms[k]]
                out['ms'][leaf] = {self.FORMS[f]: round(ms[k * 3 + f], 3) for f in range(3)}
        return out

    def run_steps(self, first, count, stream=None):
        rc = self.z.lib().zav1_wf_run_steps(self.h, first, count, stream)
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_run_steps (%s)' % self.z.lib().zav1_wf_last_error().decode())

    def run_plain(self, first, count, stream=None):
        rc = self.z.lib().zav1_wf_run_plain(self.h, first, count, stream)
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_run_plain (%s)' % self.z.lib().zav1_wf_last_error().decode())

    def recon(self):
        out = np.zeros((self.n * self.H, self.W), self.dt)
        self.z.check(self.z.lib().zav1_memcpy_d2h(out.ctypes.data, self.info.d_recon, out.nbytes, None), 'recon')
        return out.reshape(self.n, self.H, self.W)

    def units(self):
        n = self.info.n_leaves
        out = np.zeros(n, MD_UNIT)
        self.z.check(self.z.lib().zav1_memcpy_d2h(out.ctypes.data, self.info.d_units, out.nbytes, None), 'units')
        return out

    def raw_decisions(self):
        """The device's decision records in storage order (one per leaf)."""
        d = np.zeros(self.info.n_leaves, DECISION)
        self.z.check(self.z.lib().zav1_memcpy_d2h(d.ctypes.data, self.info.d_decisions, d.nbytes, None), 'dec')
        return d

    def set_raw_decisions(self, d):
        d = np.ascontiguousarray(d, DECISION)
        assert d.shape == (self.info.n_leaves,)
        self.z.check(self.z.lib().zav1_memcpy_h2d(self.info.d_decisions, d.ctypes.data, d.nbytes, None), 'dec')

    def decisions_cells(self):
        """Per 8x8 cell: (n_images, height / 8, width / 8) records, each leaf's
        at its top-left cell (zeros elsewhere): the layout for partition maps."""
        n = self.info.n_leaves
        d = np.zeros(n, DECISION)
        self.z.check(self.z.lib().zav1_memcpy_d2h(d.ctypes.data, self.info.d_decisions, d.nbytes, None), 'dec')
        u = self.units()
        grid = np.zeros((self.n, self.H // 8, self.W // 8), DECISION)
        img = u['y'] // self.H
        grid[img, (u['y'] - img * self.H) // 8, u['x'] // 8] = d
        return grid

    def decisions(self):
        """Per leaf grid position: (n_images, height / leaf, width / leaf) decision records."""
        n = self.info.n_leaves
        d = np.zeros(n, DECISION)
        self.z.check(self.z.lib().zav1_memcpy_d2h(d.ctypes.data, self.info.d_decisions, d.nbytes, None), 'dec')
        u = self.units()
        L = int(self.cfg[0]['leaf'])
        grid = np.zeros((self.n, self.H // L, self.W // L), DECISION)
        img = u['y'] // self.H
        grid[img, (u['y'] - img * self.H) // L, u['x'] // L] = d
        return grid

    def write_stills(self, first=0, n=None, threads=0, stream=None, cap=None):
        """The AV1 streams of images [first, first + n) of a plan made with
        keep_levels, written on `threads` host threads (0: all): a list of
        bytes objects.  `cap` is the first output buffer's size (default half
        a byte per sample); a short buffer costs a second write."""
        n = self.n - first if n is None else n
        sizes = (ctypes.c_size_t * n)()
        L = self.z.lib()
        cap = max(1 << 16, n * self.W * self.H // 2) if cap is None else cap
        out = np.empty(cap, np.uint8)
        rc = L.zav1_wf_write_stills(self.h, first, n, threads, out.ctypes.data, out.size, sizes, stream)
        if rc == self.z.ZAV1_ENOMEM:
            out = np.empty(max(1, sum(sizes)), np.uint8)
            rc = L.zav1_wf_write_stills(self.h, first, n, threads, out.ctypes.data, out.size, sizes, stream)
        if rc:
            raise self.z.Zav1Error(rc, 'zav1_wf_write_stills (%s)' % L.zav1_wf_last_error().decode())
        res, o = [], 0
        for k in range(n):
            res.append(out[o:o + sizes[k]].tobytes())
            o += sizes[k]
        return res

    def close(self):
        if self.h:
            self.z.lib().zav1_wf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter teardown
            pass
