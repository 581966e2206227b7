"""The full-RD transform unit composed from the oracle's pieces (the checker
for zav1_txunit_rd_batch): tx_unit_inner with do_rdoq / quantize_b and
RateMode::Exact (rust/crates/svtav1-encoder/src/leaf_funnel/
tx_pipeline.rs:268-640)."""
import numpy as np

# TX_SCALE_TAB (enc/quant.rs:238): (samples > 256) + (samples > 1024)
TX_SCALE = [0, 0, 0, 1, 2, 0, 0, 0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 1, 1]


def tx_class(tx_type):
    return 0 if tx_type < 10 else (1 if tx_type & 1 else 2)


def scan_class(tx_type):
    return 0 if tx_type < 10 else (2 if tx_type & 1 else 1)


def eob_of(q, scan):
    nz = np.nonzero(np.asarray(q).ravel()[scan])[0]
    return int(nz[-1]) + 1 if nz.size else 0


def rd_unit(oracle, src_blk, pred, tx_size, tx_type, qidx, bd, tables, *, rdoq=None, rdmult=0, tsc=0, dsc=0,
            txt_rate=0, plane=0, spatial=False, crop=None, sharp=0):
    """One unit: src_blk / pred h x w integer arrays.  rdoq None = quantize_b,
    else (cut_off_num, cut_off_denum) after quantize_fp.  Returns
    (eob, rate, dist, cul, q packed, recon)."""
    w, h = oracle.tx_dims(tx_size)
    res = src_blk.astype(np.int64) - pred.astype(np.int64)
    coef = oracle.fwd_txfm2d(res.astype(np.int32), tx_size, tx_type)
    if coef is None:
        return None
    pw, ph = min(w, 32), min(h, 32)
    packed = np.ascontiguousarray(coef[:ph, :pw])
    row = oracle.quant_row(qidx, bd)
    scan = oracle.scan(tx_size, scan_class(tx_type))
    ls = TX_SCALE[tx_size]
    txb, eobc = tables
    _, txs_ctx, ems = oracle.rdoq_size_ctx(tx_size)
    if rdoq is None:
        q, dq = oracle.quantize('b', packed, row, ls, bd)
        eob = eob_of(q, scan)
    else:
        q, dq = oracle.quantize('fp', packed, row, ls, bd)
        eob = eob_of(q, scan)
        if eob:
            q, dq, eob = oracle.optimize_b(packed.ravel(), q, dq, eob, scan, [row.dequant[0], row.dequant[1]],
                                           txb[txs_ctx * 2 + plane], eobc[ems, plane], rdmult, tx_size,
                                           tx_class(tx_type), tsc, dsc, sharp, rdoq[0], rdoq[1])
    c = txb[txs_ctx * 2 + plane]
    if eob == 0:
        rate = int(c[tsc * 2 + 1])
    else:
        rate = oracle.cost_coeffs_txb(np.asarray(q).ravel(), eob, scan, c, eobc[ems, plane], tx_size,
                                      tx_class(tx_type), plane, tsc, dsc, txt_rate)
    qa = np.asarray(q).reshape(ph, pw)
    lsum = min(int(np.abs(qa).sum()), 63)
    cul = lsum | (1 << 6 if qa[0, 0] < 0 else (2 << 6 if qa[0, 0] > 0 else 0))
    full = np.zeros((h, w), np.int32)
    full[:ph, :pw] = np.asarray(dq).reshape(ph, pw)
    inv = oracle.inv_txfm2d(full, tx_size, tx_type, bd) if eob else np.zeros((h, w), np.int32)
    rec = np.clip(pred.astype(np.int64) + inv, 0, (1 << bd) - 1)
    if spatial:
        cw, ch = crop if crop is not None else (w, h)
        dist = int(((src_blk.astype(np.int64) - rec)[:ch, :cw] ** 2).sum()) << 4
    else:
        d = int(((packed.astype(np.int64) - np.asarray(dq).reshape(ph, pw)) ** 2).sum())
        d += int((coef.astype(np.int64) ** 2).sum() - (packed.astype(np.int64) ** 2).sum())
        sh = (1 - ls) * 2
        dist = d >> sh if sh >= 0 else d << -sh
    return eob, rate, dist, cul, qa, rec
