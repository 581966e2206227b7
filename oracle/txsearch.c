/*
 * Oracle: scan orders, eob, and the fused transform-unit evaluation used by
 * the tx-type search (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h).
 *
 *   scan()              rust/crates/svtav1-entropy/src/scan_tables.rs:1640
 *                       (C-extracted tables; restated as the diagonal /
 *                       row / column rule they follow, pinned by digests)
 *   eob_from_qcoeff     rust/crates/svtav1-encoder/src/quant.rs:318
 *   tx_unit_inner       rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268
 *                       residual -> fwd -> 64-fold (+three_quad, :335-365)
 *                       -> quantize_fp -> freq distortion (:567-577)
 *   predict_dc          rust/crates/svtav1-dsp/src/intra_pred.rs:18
 */
#include <stdlib.h>
#include <string.h>

#include "zav1_oracle.h"

/* TX_TYPE_TO_SCAN_INDEX (scan_tables.rs:12): 0 diagonal, 1 row, 2 column. */
int or_scan_class(int tx_type) {
  static const int T[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 1, 2, 1, 2};
  return T[tx_type];
}

/* Scan of the (packed, <= 32x32) coefficient block of tx_size. */
int or_scan(int tx_size, int scan_class, uint16_t *scan) {
  int w, h;
  or_tx_size_dims(tx_size, &w, &h);
  if (w > 32) w = 32;
  if (h > 32) h = 32;
  int k = 0;
  if (scan_class == 1) {
    for (int i = 0; i < w * h; i++) scan[k++] = (uint16_t)i;
  } else if (scan_class == 2) {
    for (int c = 0; c < w; c++)
      for (int r = 0; r < h; r++) scan[k++] = (uint16_t)(r * w + c);
  } else {
    for (int d = 0; d < w + h - 1; d++) {
      /* square: alternate direction (even diagonals run up-right);
       * wide blocks run up-right, tall blocks down-left. */
      int up = (w == h) ? ((d & 1) == 0) : (w > h);
      for (int i = 0; i < h; i++) {
        int r = up ? (h - 1 - i) : i, c = d - r;
        if (c >= 0 && c < w) scan[k++] = (uint16_t)(r * w + c);
      }
    }
  }
  return k;
}

int or_eob(const int32_t *qcoeff, const uint16_t *scan, int n) {
  for (int i = n - 1; i >= 0; i--)
    if (qcoeff[scan[i]] != 0) return i + 1;
  return 0;
}

/* tx_pipeline.rs:642 energy_region */
static uint64_t energy(const int32_t *c, int stride, int w, int h) {
  uint64_t e = 0;
  for (int r = 0; r < h; r++)
    for (int x = 0; x < w; x++) e += (uint64_t)((int64_t)c[r * stride + x] * c[r * stride + x]);
  return e;
}

int or_tx_unit(const int32_t *residual, int tx_size, int tx_type, int qindex, int bd, int use_b,
               int32_t *qcoeff_out, uint64_t *dist, int *eob, uint32_t *level_sum) {
  int w, h;
  or_tx_size_dims(tx_size, &w, &h);
  int32_t *coeff = (int32_t *)malloc(sizeof(int32_t) * w * h);
  if (!or_fwd_txfm2d(residual, coeff, w, tx_size, tx_type)) { free(coeff); return 0; }
  int pw = w < 32 ? w : 32, ph = h < 32 ? h : 32;
  uint64_t tq = 0;
  if (w == 64 && h == 64) tq = energy(coeff + 32, 64, 32, 32) + energy(coeff + 32 * 64, 64, 64, 32);
  else if (w == 64) tq = energy(coeff + 32, 64, 32, h < 32 ? h : 32);
  else if (h == 64) tq = energy(coeff + 32 * w, w, w, h - 32);
  int32_t packed[1024], q[1024], dq[1024];
  for (int r = 0; r < ph; r++)
    for (int c = 0; c < pw; c++) packed[r * pw + c] = coeff[r * w + c];
  free(coeff);
  or_quant_row row;
  or_build_quant_row(qindex, bd, &row);
  int ls = or_tx_scale(tx_size);
  if (use_b) or_quantize_b(packed, pw * ph, &row, ls, bd, q, dq);
  else or_quantize_fp(packed, pw * ph, &row, ls, bd, q, dq);
  uint16_t scan[1024];
  int n = or_scan(tx_size, or_scan_class(tx_type), scan);
  int e = or_eob(q, scan, n);
  uint64_t d = e > 0 ? or_sse_i32(packed, dq, pw * ph) : or_sq_sum_i32(packed, pw * ph);
  d += tq;
  int shift = (1 - ls) * 2;
  d = shift < 0 ? (d << (-shift)) : (d >> shift);
  uint32_t ls_sum = 0;
  for (int i = 0; i < pw * ph; i++) ls_sum += (uint32_t)(q[i] < 0 ? -q[i] : q[i]);
  if (qcoeff_out) memcpy(qcoeff_out, q, sizeof(int32_t) * pw * ph);
  *dist = d;
  *eob = e;
  *level_sum = ls_sum;
  return 1;
}

/* intra_pred.rs:18 predict_dc, sample-generic (8- or 16-bit neighbours). */
int32_t or_dc_value(const int32_t *above, const int32_t *left, int w, int h, int has_above,
                    int has_left, int bd) {
  uint32_t sum = 0;
  if (has_above && has_left) {
    for (int i = 0; i < w; i++) sum += (uint32_t)above[i];
    for (int i = 0; i < h; i++) sum += (uint32_t)left[i];
    uint32_t cnt = (uint32_t)(w + h);
    return (int32_t)((sum + cnt / 2) / cnt);
  }
  if (has_above) {
    for (int i = 0; i < w; i++) sum += (uint32_t)above[i];
    return (int32_t)((sum + (uint32_t)w / 2) / (uint32_t)w);
  }
  if (has_left) {
    for (int i = 0; i < h; i++) sum += (uint32_t)left[i];
    return (int32_t)((sum + (uint32_t)h / 2) / (uint32_t)h);
  }
  return 1 << (bd - 1);
}

/* Reference for the batched tx-type search over a u16 plane: every block
 * (x0, y0) of tx_size, DC prediction from the source row above / column
 * left, every type in `types`.  Outputs indexed [block * n_types + k]. */
void or_txsearch_plane(const uint16_t *plane, int stride, int width, int height, int bd, int tx_size,
                       const int32_t *xy, int n_blocks, const uint8_t *types, int n_types, int qindex,
                       uint64_t *dist, int32_t *eob, uint32_t *lsum) {
  int w, h;
  or_tx_size_dims(tx_size, &w, &h);
#pragma omp parallel for schedule(dynamic, 16)
  for (int b = 0; b < n_blocks; b++) {
    int x0 = xy[2 * b], y0 = xy[2 * b + 1];
    int32_t above[64], left[64], res[4096];
    for (int i = 0; i < w; i++) above[i] = y0 > 0 ? plane[(y0 - 1) * stride + x0 + i] : 0;
    for (int i = 0; i < h; i++) left[i] = x0 > 0 ? plane[(y0 + i) * stride + x0 - 1] : 0;
    int32_t dc = or_dc_value(above, left, w, h, y0 > 0, x0 > 0, bd);
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) res[r * w + c] = (int32_t)plane[(y0 + r) * stride + x0 + c] - dc;
    for (int k = 0; k < n_types; k++) {
      int e = 0;
      uint32_t s = 0;
      uint64_t d = 0;
      or_tx_unit(res, tx_size, types[k], qindex, bd, 0, NULL, &d, &e, &s);
      dist[b * n_types + k] = d;
      eob[b * n_types + k] = e;
      lsum[b * n_types + k] = s;
    }
  }
  (void)width;
  (void)height;
}

/* tx_unit_inner (lf/tx_pipeline.rs:268) and its u16 twin tx_unit_hbd
 * (:854) for one transform unit, without the RDOQ trellis, quantization
 * matrices, noise normalization, tx-bias / psy terms and coefficient rate:
 *   residual_i32 (dsp/residual.rs:26): src - pred
 *   forward 2D, 64-point fold + three_quad energy (:335-365)
 *   quantize: use_b ? quantize_b : quantize_fp (the RDOQ path's quantizer)
 *   eob in scan order; cul level compute_cul_level (:130)
 *   recon when need_recon or spatial (:432-470 / hbd :1004-1030):
 *     eob > 0 ? clamp(pred + inverse 2D of the unfolded dqcoeff, 0, 2^bd - 1)
 *             : pred
 *   dist: spatial ? SSE(src, recon) over crop_w x crop_h << 4
 *         (variance.rs:60 / hbd.rs:936)
 *       : (sse_i32(packed, dq) or sq_sum_i32(packed)) + three_quad,
 *         shifted by (1 - log_scale) * 2 (:567-577).
 * src: u8 (hbd == 0) or u16 plane; pred: u16 w x h raster.  Outputs:
 * qcoeff (packed pw x ph, may be NULL), recon (u16 w x h, NULL when not
 * reconstructed). */
int or_tx_unit_full(const void *src, int src_stride, int hbd, const uint16_t *pred, int tx_size, int tx_type,
                    int qindex, int bd, int use_b, int need_recon, int spatial, int crop_w, int crop_h,
                    int32_t *qcoeff_out, uint16_t *recon_out, uint64_t *dist, int *eob, uint8_t *cul) {
  int w, h;
  or_tx_size_dims(tx_size, &w, &h);
  const int n = w * h;
  int32_t *res = (int32_t *)malloc(sizeof(int32_t) * n);
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      const long o = (long)r * src_stride + c;
      const int32_t s = hbd ? ((const uint16_t *)src)[o] : ((const uint8_t *)src)[o];
      res[r * w + c] = s - (int32_t)pred[r * w + c];
    }
  int32_t *coeff = (int32_t *)malloc(sizeof(int32_t) * n);
  if (!or_fwd_txfm2d(res, coeff, w, tx_size, tx_type)) {
    free(res);
    free(coeff);
    return 0;
  }
  const int pw = w < 32 ? w : 32, ph = h < 32 ? h : 32;
  uint64_t tq = 0;
  if (w == 64 && h == 64) tq = energy(coeff + 32, 64, 32, 32) + energy(coeff + 32 * 64, 64, 64, 32);
  else if (w == 64) tq = energy(coeff + 32, 64, 32, h < 32 ? h : 32);
  else if (h == 64) tq = energy(coeff + 32 * w, w, w, h - 32);
  int32_t packed[1024], q[1024], dq[1024];
  for (int r = 0; r < ph; r++)
    for (int c = 0; c < pw; c++) packed[r * pw + c] = coeff[r * w + c];
  or_quant_row row;
  or_build_quant_row(qindex, bd, &row);
  const int ls = or_tx_scale(tx_size);
  if (use_b) or_quantize_b(packed, pw * ph, &row, ls, bd, q, dq);
  else or_quantize_fp(packed, pw * ph, &row, ls, bd, q, dq);
  uint16_t scan[1024];
  const int ns = or_scan(tx_size, or_scan_class(tx_type), scan);
  const int e = or_eob(q, scan, ns);
  /* compute_cul_level: |q| in scan order up to eob, capped at 63, dc sign */
  uint32_t cl = 0;
  for (int i = 0; i < e; i++) {
    const int32_t v = q[scan[i]];
    cl += (uint32_t)(v < 0 ? -v : v);
    if (cl >= 63) break;
  }
  if (cl > 63) cl = 63;
  const int32_t dc = e > 0 ? q[0] : 0;
  if (dc < 0) cl |= 1u << 6;
  else if (dc > 0) cl += 2u << 6;
  *cul = (uint8_t)cl;
  *eob = e;
  if (qcoeff_out) memcpy(qcoeff_out, q, sizeof(int32_t) * pw * ph);

  uint16_t *rec = NULL;
  if (need_recon || spatial) {
    rec = recon_out ? recon_out : (uint16_t *)malloc(sizeof(uint16_t) * n);
    if (e > 0) {
      int32_t *dq_full = (int32_t *)calloc((size_t)n, sizeof(int32_t));
      for (int r = 0; r < ph; r++)
        for (int c = 0; c < pw; c++) dq_full[r * w + c] = dq[r * pw + c];
      or_inv_txfm2d(dq_full, res, w, tx_size, tx_type, bd);
      const int32_t maxv = (1 << bd) - 1;
      for (int i = 0; i < n; i++) {
        const int32_t v = (int32_t)pred[i] + res[i];
        rec[i] = (uint16_t)(v < 0 ? 0 : (v > maxv ? maxv : v));
      }
      free(dq_full);
    } else {
      memcpy(rec, pred, sizeof(uint16_t) * n);
    }
  }
  if (spatial) {
    uint64_t s = 0;
    for (int r = 0; r < crop_h; r++)
      for (int c = 0; c < crop_w; c++) {
        const long o = (long)r * src_stride + c;
        const int64_t sv = hbd ? ((const uint16_t *)src)[o] : ((const uint8_t *)src)[o];
        const int64_t d = sv - rec[r * w + c];
        s += (uint64_t)(d * d);
      }
    *dist = s << 4;
  } else {
    uint64_t d = e > 0 ? or_sse_i32(packed, dq, pw * ph) : or_sq_sum_i32(packed, pw * ph);
    d += tq;
    const int shift = (1 - ls) * 2;
    *dist = shift < 0 ? (d << (-shift)) : (d >> shift);
  }
  if (rec && rec != recon_out) free(rec);
  free(res);
  free(coeff);
  return 1;
}
