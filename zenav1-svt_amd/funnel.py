"""Host driver of the intra mode-decision funnel's first two stages on the
device: MDS0 (predict every candidate mode of every unit, Hadamard SATD,
fast cost, survivor selection) feeding MDS1 (transform units of the
survivors), without the candidate tables leaving HBM between stages.

Reference path (rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs):
candidate prediction predict_unit (predict.rs:49) + hadamard_satd
(predict.rs:143) in the fast loop (mod.rs:2060-2185), fast cost rdcost
(rate_tables.rs:1033), survivor order by the exchange sort
(rate_tables.rs:108, mod.rs:2994), MDS1 = DCT_DCT with quantize_b and
frequency-domain SSE (mod.rs:40-44, tx_unit_inner tx_pipeline.rs:268).

Here every unit of a frame-sized grid is evaluated in one batch against a
given neighbour canvas (`recon`): the batched form of the fast loop for a
set of units whose neighbours are final (a wavefront, or a pre-analysis
pass).  Mode rates come from the caller (context-dependent table lookups on
the host); the device adds them to the SATD term.
"""
import numpy as np

# AV1 transform sizes in TxSize order (w, h).
TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
           (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]

# Candidate list of one unit: (mode, angle delta, filter-intra mode); 5 = no
# filter intra.  The 13 regular luma modes in generation order.
REGULAR_MODES = [(m, 0, 5) for m in range(13)]


def angle_candidates():
    """13 regular modes plus the directional angle deltas +-1..3 (61)."""
    c = list(REGULAR_MODES)
    for m in range(1, 9):
        for d in (-3, -2, -1, 1, 2, 3):
            c.append((m, d, 5))
    return c


def tx_size_of(w, h):
    return TX_DIMS.index((w, h))


def unit_grid(width, height, w, h):
    ys, xs = np.mgrid[0:height - h + 1:h, 0:width - w + 1:w]
    return np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.int32)


def raster_availability(x, y, w, h, width):
    """Neighbour sample counts of a unit of a uniform grid coded in raster
    order: top and top-right from the row above (top-right limited by the
    frame), left from the unit before, bottom-left not yet coded."""
    nt = w if y > 0 else 0
    ntr = max(0, min(w, width - (x + w))) if y > 0 else 0
    nl = h if x > 0 else 0
    return nt, ntr, nl, 0


class IntraFunnel:
    """MDS0 -> MDS1 over every w x h unit of a plane, device resident.

    fused=True (default): one fused launch predicts every candidate on chip
    and returns its SATD (zav1_md_fast_satd); after the selection only the
    k survivors per unit are predicted again, for the transform units.
    fused=False: every candidate's prediction is written to HBM
    (zav1_intra_pred_batch) and read back by zav1_pred_satd_batch."""

    def __init__(self, zav1, src, recon, bd, w, h, qindex, k=3, cands=None, lam=40000, rates=None, types=(0,),
                 flags=None, edge_filter=1, fused=True):
        self.zav1 = zav1
        self.fused = fused
        self.bd, self.w, self.h, self.qindex, self.k = bd, w, h, qindex, k
        self.tx_size = tx_size_of(w, h)
        dt = np.uint8 if bd == 8 else np.uint16
        self.src = np.ascontiguousarray(src, dtype=dt)
        self.recon = np.ascontiguousarray(recon, dtype=dt)
        self.ph, self.pw = self.src.shape
        self.cands = list(REGULAR_MODES if cands is None else cands)
        nc = self.n_cands = len(self.cands)
        if not 1 <= k <= nc <= 64:
            raise ValueError('need 1 <= k <= n_cands <= 64')
        self.types = np.asarray(types, np.uint8)
        self.flags = zav1.TXU_QUANT_B if flags is None else flags
        self.lam = int(lam)
        self.xy = unit_grid(self.pw, self.ph, w, h)
        nu = self.n_units = self.xy.shape[0]
        n = w * h
        ij = np.zeros(nu * nc, zav1.INTRA_JOB)
        sj = np.zeros(nu * nc, zav1.TXUNIT_JOB)
        for u, (x, y) in enumerate(self.xy.tolist()):
            nt, ntr, nl, nbl = raster_availability(x, y, w, h, self.pw)
            for m, (mode, delta, fi) in enumerate(self.cands):
                j = u * nc + m
                ij[j] = (x, y, w, h, mode, delta, fi, edge_filter, 0, 0, nt, ntr, nl, nbl, j * n, 0)
                sj[j] = (x, y, j * n, 0, w, h, 0)
        self.intra_jobs, self.satd_jobs = ij, sj
        mu = np.zeros(nu, zav1.MD_UNIT)
        for u, (x, y) in enumerate(self.xy.tolist()):
            mu[u] = (x, y) + raster_availability(x, y, w, h, self.pw) + (edge_filter, 0, 0, 0)
        self.md_units = mu
        self.md_cands = np.array([(m, d, fi, 0) for (m, d, fi) in self.cands], zav1.MD_CAND)
        if rates is None:
            rates = np.zeros((nu, nc), np.uint32)
        self.rates = np.ascontiguousarray(np.broadcast_to(rates, (nu, nc)), dtype=np.uint32)
        A = zav1.DeviceArray
        self.d_src = A.from_host(self.src)
        self.d_recon = A.from_host(self.recon)
        self.d_ij = A.from_host(ij.view(np.uint8))
        self.d_sj = A.from_host(sj.view(np.uint8))
        self.d_xy = A.from_host(self.xy)
        self.d_rate = A.from_host(self.rates)
        self.d_units = A.from_host(mu.view(np.uint8))
        self.d_pred = A(nu * (k if fused else nc) * n, np.uint16)
        self.d_satd = A(nu * nc, np.uint32)
        self.d_sel = A(nu * k, np.uint8)
        self.d_cost = A(nu * k, np.uint64)
        nt = nu * k * len(self.types)
        self.n_tx = nt
        self.d_tj = A(nt * zav1.TXUNIT_JOB.itemsize, np.uint8)
        self.d_dist = A(nt, np.uint64)
        self.d_eob = A(nt, np.int32)
        self.d_cul = A(nt, np.uint8)

    def predictions(self):
        return self.n_units * self.n_cands

    def stages(self):
        return ('satd', 'select', 'pred', 'tx') if self.fused else ('pred', 'satd', 'select', 'tx')

    def run(self, stream=None, stages=None):
        z = self.zav1
        stages = self.stages() if stages is None else stages
        if self.fused:
            for st in stages:
                if st == 'satd':
                    z.md_fast_satd(self.d_src, self.d_recon, self.pw, self.pw, self.ph, self.bd, self.w, self.h,
                                   self.d_units, self.n_units, self.md_cands, self.d_satd, stream)
                elif st == 'select':
                    z.md_fast_select(self.d_satd, self.d_rate, self.lam, self.n_units, self.n_cands, self.k,
                                     self.d_units, self.w, self.h, self.types, self.d_sel, self.d_cost, self.d_tj,
                                     stream, xy_stride=self.zav1.MD_UNIT.itemsize // 4, survivors_only=True)
                elif st == 'pred':
                    z.md_predict_survivors(self.d_recon, self.pw, self.pw, self.ph, self.bd, self.w, self.h,
                                           self.d_units, self.n_units, self.md_cands, self.d_sel, self.k,
                                           self.d_pred, stream)
                elif st == 'tx':
                    z.txunit_batch(self.d_src, self.pw, self.pw, self.ph, self.bd, self.tx_size, self.d_pred,
                                   self.d_tj, self.n_tx, self.qindex, self.flags, self.d_dist, self.d_eob,
                                   self.d_cul, None, None, stream)
            return
        if 'pred' in stages:
            z.intra_pred_batch(self.d_recon, self.pw, self.bd, self.d_ij, self.predictions(), self.d_pred, stream)
        if 'satd' in stages:
            z.pred_satd_batch(self.d_src, self.pw, self.pw, self.ph, self.bd, self.w, self.h, self.d_pred, self.d_sj,
                              self.predictions(), self.d_satd, stream)
        if 'select' in stages:
            z.md_fast_select(self.d_satd, self.d_rate, self.lam, self.n_units, self.n_cands, self.k, self.d_xy,
                             self.w, self.h, self.types, self.d_sel, self.d_cost, self.d_tj, stream)
        if 'tx' in stages:
            z.txunit_batch(self.d_src, self.pw, self.pw, self.ph, self.bd, self.tx_size, self.d_pred, self.d_tj,
                           self.n_tx, self.qindex, self.flags, self.d_dist, self.d_eob, self.d_cul, None, None, stream)

    def results(self):
        nu, k, nty = self.n_units, self.k, len(self.types)
        return dict(satd=self.d_satd.to_host().reshape(nu, self.n_cands),
                    sel=self.d_sel.to_host().reshape(nu, k),
                    sel_cost=self.d_cost.to_host().reshape(nu, k),
                    jobs=self.d_tj.to_host().view(self.zav1.TXUNIT_JOB).reshape(nu, k, nty),
                    dist=self.d_dist.to_host().reshape(nu, k, nty),
                    eob=self.d_eob.to_host().reshape(nu, k, nty),
                    cul=self.d_cul.to_host().reshape(nu, k, nty))

    def survivor_pred(self, u, rank):
        """Prediction of survivor `rank` of unit u (after a full run)."""
        n = self.w * self.h
        if self.fused:
            off = (u * self.k + rank) * n
        else:
            off = (u * self.n_cands + int(self.d_sel.to_host()[u * self.k + rank])) * n
        return self.d_pred.to_host()[off:off + n].reshape(self.h, self.w)
