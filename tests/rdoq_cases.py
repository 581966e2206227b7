"""Seeded RDOQ inputs: transform-like coefficient blocks fp-quantized by the
oracle, with their scans and eobs (shared by the CPU and GPU tests)."""
import numpy as np

TXW = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64]
TXH = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16]
TX_SCALE = [0, 0, 0, 1, 2, 0, 0, 0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 1, 1]
SCAN_CLASS = {0: 0, 1: 2, 2: 1}  # tx_class -> scan class (vertical class scans rows)


def coded_dims(tx_size):
    return min(TXW[tx_size], 32), min(TXH[tx_size], 32)


def spectrum(rng, n_blocks, w, h, amp):
    """Coefficients decaying away from DC, random signs, a few outliers."""
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    env = amp / (1.0 + 0.35 * (r + c))
    mag = rng.exponential(1.0, (n_blocks, h, w)) * env
    mag[rng.random((n_blocks, h, w)) < 0.01] *= 12
    sgn = np.where(rng.random((n_blocks, h, w)) < 0.5, -1, 1)
    return (np.round(mag) * sgn).astype(np.int32).reshape(n_blocks, w * h)


def blocks(oracle, tx_size, tx_class, n_blocks, seed, qindex, bd, amp=None):
    """(tcoeff, qcoeff, dqcoeff, eob, scan) with fp quantization at qindex."""
    rng = np.random.default_rng(seed)
    w, h = coded_dims(tx_size)
    amp = amp or (600 if bd == 8 else 2400) * (1 << TX_SCALE[tx_size])
    t = spectrum(rng, n_blocks, w, h, amp)
    row = oracle.quant_row(qindex, bd)
    scan = oracle.scan(tx_size, SCAN_CLASS[tx_class])
    q = np.zeros_like(t)
    dq = np.zeros_like(t)
    eob = np.zeros(n_blocks, np.int32)
    for b in range(n_blocks):
        q[b], dq[b] = oracle.quantize('fp', t[b], row, TX_SCALE[tx_size], bd)
        nz = np.nonzero(q[b][scan])[0]
        eob[b] = nz[-1] + 1 if nz.size else 0
    return t, q, dq, eob, scan, [row.dequant[0], row.dequant[1]]
