"""Host driver of the batched tx-type search (candidate RD over transform
sizes and types) — the caller side of the fused transform-unit kernel.

For a frame, every one of the 19 transform sizes tiles the frame, and every
block is evaluated with every transform type the encoder may signal for that
size (AV1: all 16 up to 16 points per side; DCT/IDTX and the 1D-DCT/identity
mixes when a side is 32; DCT only when a side is 64).  This is the
exhaustive form of the MDS3 tx-type search
(rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs:7140-7330) over every tx
size at once; one launch per (size, qindex) covers the whole frame.
"""
import numpy as np

TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
           (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def legal_types(tx_size):
    w, h = TX_DIMS[tx_size]
    if max(w, h) == 64:
        return [0]
    if max(w, h) == 32:
        return [0, 9, 10, 11]  # DCT_DCT, IDTX, V_DCT, H_DCT
    return list(range(16))


def tile_xy(width, height, tx_size):
    w, h = TX_DIMS[tx_size]
    ys, xs = np.mgrid[0:height - h + 1:h, 0:width - w + 1:w]
    return np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.int32)


class TxSearchSweep:
    """Device-resident tx-type search over one frame at several qindices."""

    def __init__(self, zav1, plane, bd, qindices, sizes=None):
        self.zav1 = zav1
        self.plane = np.ascontiguousarray(plane, dtype=np.uint16)
        self.h, self.w = self.plane.shape
        self.bd = bd
        self.qindices = list(qindices)
        self.sizes = list(range(19)) if sizes is None else list(sizes)
        self.d_plane = zav1.DeviceArray.from_host(self.plane)
        self.jobs = []
        for ts in self.sizes:
            xy = tile_xy(self.w, self.h, ts)
            types = np.array(legal_types(ts), np.uint8)
            d_xy = zav1.DeviceArray.from_host(xy)
            per_q = []
            for _ in self.qindices:
                n = xy.shape[0] * types.size
                per_q.append((zav1.DeviceArray(n, np.uint64), zav1.DeviceArray(n, np.int32),
                              zav1.DeviceArray(n, np.uint32)))
            self.jobs.append((ts, xy, types, d_xy, per_q))

    def candidates_per_step(self):
        return sum(xy.shape[0] * t.size for _, xy, t, _, _ in self.jobs) * len(self.qindices)

    def coeffs_per_step(self):
        tot = 0
        for ts, xy, t, _, _ in self.jobs:
            w, h = TX_DIMS[ts]
            tot += xy.shape[0] * t.size * w * h
        return tot * len(self.qindices)

    def run(self, stream=None, sizes=None):
        for ts, xy, types, d_xy, per_q in self.jobs:
            if sizes is not None and ts not in sizes:
                continue
            for qi, (dd, de, dl) in zip(self.qindices, per_q):
                self.zav1.txsearch_batch(self.d_plane, self.w, self.bd, ts, d_xy, xy.shape[0], types, qi, dd, de,
                                         dl, stream)

    def results(self, tx_size, q_pos=0):
        for ts, xy, types, d_xy, per_q in self.jobs:
            if ts == tx_size:
                dd, de, dl = per_q[q_pos]
                return xy, types, dd.to_host(), de.to_host(), dl.to_host()
        raise KeyError(tx_size)
