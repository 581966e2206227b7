"""Device-resident loop-filter stage of one 4:2:0 still: deblocking of all
three planes, the CDEF direction + 64-strength SSE tables per 64x64 filter
block, and the Wiener / self-guided statistics and unfiltered SSE of every
restoration unit.

The order follows the reference encoder's still-image path
(rust/crates/svtav1-encoder/src/pipeline.rs:144: deblock -> CDEF search ->
restoration search): CDEF searches the deblocked frame, and restoration
statistics are taken on the deblocked frame here (the workload stops at the
statistics; no CDEF strength is applied between the stages).
"""
import numpy as np


class LoopFilterStage:
    """Holds source / reconstruction planes, geometry and every output in HBM.

    planes: dict with 'src' and 'rec' -> (y, u, v) numpy arrays (u8 at bd 8,
    u16 at bd 10); geom: dict of mi arrays from synth.partition_geometry.  An
    all-intra still has no inter-skip, so deblocking sees no skip map; the
    geometry's 'skip' (signaled skip per mi) drives the CDEF search list.
    """

    def __init__(self, zav1, planes, geom, bd, qindex, levels, sharpness=0, cdef_sub=1, ru_luma=128,
                 ru_chroma=64, sgr=True):
        self.z = zav1
        self.bd = bd
        self.qindex = qindex
        self.levels = levels
        self.sharpness = sharpness
        self.cdef_sub = cdef_sub
        self.sgr = sgr
        y = planes['src'][0]
        self.h, self.w = y.shape
        self.ch, self.cw = planes['src'][1].shape
        D = zav1.DeviceArray
        self.src = [D.from_host(p) for p in planes['src']]
        self.rec = [D.from_host(p) for p in planes['rec']]
        self.dgd = [D(p.shape, p.dtype) for p in planes['rec']]
        mi_rows, mi_cols = geom['bw'].shape
        self.mi_rows, self.mi_cols = mi_rows, mi_cols
        self.geom_dev = {k: D.from_host(v) for k, v in geom.items() if k != 'skip'}
        self.geom = zav1.deblock_geom(mi_cols, mi_rows, self.w, self.h, self.geom_dev)
        self.cdef_skip = D.from_host(geom['skip']) if 'skip' in geom else None
        nfb = ((self.w + 63) // 64) * ((self.h + 63) // 64)
        self.nfb = nfb
        self.dirvar = D((nfb, 64, 2), np.int32)
        self.mse_y = D((nfb, 64), np.uint64)
        self.mse_uv = D((nfb, 64), np.uint64)
        self.ru = [ru_luma, ru_chroma, ru_chroma]
        self.grid = [zav1.rest_unit_grid(ru_luma, self.w, self.h)] + \
            [zav1.rest_unit_grid(ru_chroma, self.cw, self.ch)] * 2
        self.win = [7, 5, 5]
        self.M, self.H, self.sse, self.sgrs = [], [], [], []
        for p in range(3):
            n = self.grid[p][0] * self.grid[p][1]
            w2 = self.win[p] ** 2
            self.M.append(D((n, w2), np.int64))
            self.H.append(D((n, w2, w2), np.int64))
            self.sse.append(D((n,), np.uint64))
            self.sgrs.append(D((n, 16, 5), np.int64))
        self.nbytes = 2 * sum(a.nbytes for a in self.src)  # src + rec, read once

    def pixels(self):
        return self.w * self.h

    def _dims(self, p):
        return (self.w, self.h, 0) if p == 0 else (self.cw, self.ch, 1)

    def copy_rec(self, stream=None):
        """Fresh deblocking input: dgd <- rec (deblocking works in place)."""
        L = self.z.lib()
        for p in range(3):
            self.z.check(L.zav1_memcpy_d2d(self.dgd[p].ptr, self.rec[p].ptr, self.rec[p].nbytes, stream),
                         'memcpy_d2d')

    def deblock(self, stream=None):
        lv = [(self.levels[0], self.levels[1]), (self.levels[2], self.levels[2]), (self.levels[3], self.levels[3])]
        for p in range(3):
            pw, ph, ss = self._dims(p)
            self.z.deblock_plane(self.dgd[p], pw, pw, ph, self.bd, p, ss, lv[p][0], lv[p][1], self.sharpness,
                                 self.geom, stream)

    def cdef(self, stream=None):
        self.z.cdef_search(self.dgd, self.src, self.w, self.h, self.bd, self.qindex, self.cdef_sub, self.dirvar,
                           self.mse_y, self.mse_uv, stream, d_skip=self.cdef_skip, mi_cols=self.mi_cols,
                           mi_rows=self.mi_rows)

    def wiener(self, stream=None):
        for p in range(3):
            pw, ph, ss = self._dims(p)
            self.z.wiener_stats(self.dgd[p], self.src[p], pw, ph, self.bd, self.win[p], self.ru[p], ss, self.M[p],
                                self.H[p], stream)
            self.z.unit_sse(self.dgd[p], self.src[p], pw, ph, self.bd, self.ru[p], ss, self.sse[p], stream)

    def selfguided(self, stream=None):
        for p in range(3):
            pw, ph, ss = self._dims(p)
            self.z.sgr_stats(self.dgd[p], self.src[p], pw, ph, self.bd, self.ru[p], ss, self.sgrs[p], stream)

    PHASES = ('copy_rec', 'deblock', 'cdef', 'wiener', 'selfguided')

    def run(self, stream=None):
        self.copy_rec(stream)
        self.deblock(stream)
        self.cdef(stream)
        self.wiener(stream)
        if self.sgr:
            self.selfguided(stream)
