"""Host driver of the batched tx-type search (candidate RD over transform
sizes and types) — the caller side of the fused transform-unit kernel.

For a frame, every one of the 19 transform sizes tiles the frame, and every
block is evaluated with every transform type an intra block may signal at
that size: the intra ext-tx sets (rust/crates/svtav1-entropy/src/
coeff_c.rs:181-212) -- DCT_DCT only once a side reaches 32, {DCT, ADST x DCT,
ADST_ADST, IDTX} at 16x16, those and V_DCT / H_DCT at every other size.  This
is the
exhaustive form of the MDS3 tx-type search
(rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs:7140-7330) over every tx
size at once; one launch per (size, qindex) covers the whole frame.
"""
import numpy as np

TX_DIMS = [(4, 4), (8, 8), (16, 16), (32, 32), (64, 64), (4, 8), (8, 4), (8, 16), (16, 8), (16, 32),
           (32, 16), (32, 64), (64, 32), (4, 16), (16, 4), (8, 32), (32, 8), (16, 64), (64, 16)]


def legal_types(tx_size):
    """The intra ext-tx set of tx_size, in tx-type order."""
    w, h = TX_DIMS[tx_size]
    if max(w, h) >= 32:
        return [0]
    if w == 16 and h == 16:
        return [0, 1, 2, 3, 9]                 # DTT4_IDTX
    return [0, 1, 2, 3, 9, 10, 11]             # DTT4_IDTX_1DDCT


def kernel_form(tx_size):
    """Which fused kernel evaluates a size (as the library picks it): 'lane'
    (one lane per block: blocks up to 8 wide and 64 samples), 'rows' (one
    lane per row: the 16-point sizes) or 'group' (lane groups: 32- and
    64-point sizes).  Every form evaluates two quantizers per launch."""
    w, h = TX_DIMS[tx_size]
    if w * h <= 64 and w <= 8:
        return 'lane'
    if (w == 16 or h == 16) and w >= 8 and w * h <= 256:
        return 'rows'
    return 'group'


def launches(tx_size, n_q):
    """Kernel launches of one zav1_txsearch_batch_q call."""
    return (n_q + 1) // 2


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
        nq = len(self.qindices)
        for ts in self.sizes:
            xy = tile_xy(self.w, self.h, ts)
            types = np.array(legal_types(ts), np.uint8)
            d_xy = zav1.DeviceArray.from_host(xy)
            n = xy.shape[0] * types.size
            # outputs [q][block][type], all quantizers in one call
            out = (zav1.DeviceArray(nq * n, np.uint64), zav1.DeviceArray(nq * n, np.int32),
                   zav1.DeviceArray(nq * n, np.uint32))
            self.jobs.append((ts, xy, types, d_xy, out))

    def candidates_per_step(self):
        return sum(xy.shape[0] * t.size for _, xy, t, _, _ in self.jobs) * len(self.qindices)

    def coeffs_per_step(self):
        tot = 0
        for ts, xy, t, _, _ in self.jobs:
            w, h = TX_DIMS[ts]
            tot += xy.shape[0] * t.size * w * h
        return tot * len(self.qindices)

    def stream_of(self, tx_size, streams):
        """The stream a size runs on when the sweep is spread over several:
        sizes in decreasing order of work, dealt round-robin (each size
        always on the same stream, so repeated runs stay ordered per size;
        different sizes write different outputs)."""
        if not isinstance(streams, (list, tuple)):
            return streams
        order = sorted(range(19), key=lambda t: (-len(legal_types(t)), -TX_DIMS[t][0] * TX_DIMS[t][1], t))
        return streams[order.index(tx_size) % len(streams)]

    def run(self, stream=None, sizes=None):
        """One pass over the frame; `stream` may be a list of streams, over
        which the sizes are spread (see stream_of)."""
        for ts, xy, types, d_xy, (dd, de, dl) in self.jobs:
            if sizes is not None and ts not in sizes:
                continue
            self.zav1.txsearch_batch_q(self.d_plane, self.w, self.bd, ts, d_xy, xy.shape[0], types, self.qindices,
                                       dd, de, dl, self.stream_of(ts, stream))

    def results(self, tx_size, q_pos=0):
        for ts, xy, types, d_xy, (dd, de, dl) in self.jobs:
            if ts == tx_size:
                n = xy.shape[0] * types.size
                sl = slice(q_pos * n, (q_pos + 1) * n)
                return xy, types, dd.to_host()[sl], de.to_host()[sl], dl.to_host()[sl]
        raise KeyError(tx_size)
