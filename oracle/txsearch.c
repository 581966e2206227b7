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
