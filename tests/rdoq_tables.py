"""Synthetic coefficient rate tables for the RDOQ tests.

The trellis prices with tables its caller builds from the frame's CDFs
(rust/crates/svtav1-encoder/src/quant.rs:514-:632).  The default CDFs are not
part of this repository, so the tests draw per-symbol costs at random and
derive the tables' secondary entries exactly as that builder does: the base
cost deltas (entries 4..7 of base_cost), the cumulative base-range costs
lps[0..12] from four-symbol rates, and their deltas lps[13..25].  The
trellis is defined for any such table.
"""
import numpy as np

TXB_WORDS, EOB_WORDS = 944, 22
# word offsets of the TxbCosts fields (quant.rs:459)
SKIP, BASE_EOB, BASE, EOB_EXTRA, DC_SIGN, LPS = 0, 26, 38, 374, 392, 398


def symbol_costs(rng, n, scale):
    """Costs of an n-symbol alphabet in 1/512 bit: -log2 of a random
    distribution with every probability >= 4/32768."""
    p = rng.dirichlet(np.full(n, 0.7))
    p = np.maximum(p, 4 / 32768)
    p /= p.sum()
    return np.round(-np.log2(p) * 512 * scale).astype(np.int64)


def txb_costs(rng, scale=1.0):
    t = np.zeros(TXB_WORDS, np.int64)
    for ctx in range(13):
        t[SKIP + 2 * ctx:SKIP + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(4):
        t[BASE_EOB + 3 * ctx:BASE_EOB + 3 * ctx + 3] = symbol_costs(rng, 3, scale)
    for ctx in range(42):
        b = symbol_costs(rng, 4, scale)
        o = BASE + 8 * ctx
        t[o:o + 4] = b
        t[o + 4] = 0
        t[o + 5] = b[1] + 512 - b[0]
        t[o + 6] = b[2] - b[1]
        t[o + 7] = b[3] - b[2]
    for ctx in range(9):
        t[EOB_EXTRA + 2 * ctx:EOB_EXTRA + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(3):
        t[DC_SIGN + 2 * ctx:DC_SIGN + 2 * ctx + 2] = symbol_costs(rng, 2, scale)
    for ctx in range(21):
        br = symbol_costs(rng, 4, scale)
        lps = np.zeros(26, np.int64)
        prev, i = 0, 0
        while i < 12:
            for j in range(3):
                lps[i + j] = prev + br[j]
            prev += br[3]
            i += 3
        lps[12] = prev
        lps[13] = lps[0]
        for i in range(1, 13):
            lps[i + 13] = lps[i] - lps[i - 1]
        t[LPS + 26 * ctx:LPS + 26 * ctx + 26] = lps
    return t.astype(np.int32)


def eob_costs(rng, n_sym, scale=1.0):
    e = np.zeros(EOB_WORDS, np.int64)
    for ctx in range(2):
        e[11 * ctx:11 * ctx + n_sym] = symbol_costs(rng, n_sym, scale)
    return e.astype(np.int32)


def coeff_costs(seed, scale=1.0):
    """A frame's tables (zav1_coeff_costs layout): txb[10] then eob[7][2];
    eob_multi_size m codes 5 + m symbols."""
    rng = np.random.default_rng(seed)
    txb = [txb_costs(rng, scale) for _ in range(10)]
    eob = [eob_costs(rng, min(5 + m, 11), scale) for m in range(7) for _ in range(2)]
    return np.concatenate(txb + eob)


def tables_for(costs, txs_ctx, eob_ms, plane):
    t = costs[(txs_ctx * 2 + plane) * TXB_WORDS:(txs_ctx * 2 + plane + 1) * TXB_WORDS]
    o = 10 * TXB_WORDS + (eob_ms * 2 + plane) * EOB_WORDS
    return t, costs[o:o + EOB_WORDS]
