/*
 * Oracle: quantizer rows, quantizers, frequency/spatial distortion and the
 * Hadamard SATD (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "av1_qlookup_tables.h"
#include "zav1_oracle.h"

/* enc/quant.rs:238 TX_SCALE_TAB, C TxSize order */
int or_tx_scale(int tx_size) {
  static const int T[19] = {0, 0, 0, 1, 2, 0, 0, 0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 1, 1};
  return T[tx_size];
}

/* enc/quant.rs:69 invert_quant */
static void invert_quant(int32_t d, int32_t *quant, int32_t *shift) {
  uint32_t t = (uint32_t)d;
  int l = 0;
  while (t > 1) { t >>= 1; l++; }
  int64_t m = 1 + ((int64_t)1 << (16 + l)) / d;
  *quant = (int32_t)(m - (1 << 16));
  *shift = 1 << (16 - l);
}

/* enc/quant.rs:132-236 build_quant_table(_bd) with sharpness 0;
 * qzbin factor per enc/bd10.rs:52 (threshold 148 at bd8, 592 at bd10). */
/* enc/quant.rs:105 apply_quant_sharpness_factors (base qindex 31, :90):
 * sharpness s > 0 below the base (s < 0 above it) moves the zbin / rounding
 * factors by an offset of the qindex distance, clamped to [1, 256]. */
static void sharp_factors(int *qzbin, int *qround, int qindex, int s) {
  const int diff = qindex - 31;
  if (!((s > 0 && diff < 0) || (s < 0 && diff > 0))) return;
  const int ad = diff < 0 ? -diff : diff;
  const int off = s > 0 ? (2 * s > ad ? 2 * s : ad) : (-2 * s < diff ? -2 * s : diff);
  int z = s > 0 ? *qzbin - off : *qzbin + off;
  int r = s > 0 ? *qround + off : *qround - off;
  *qzbin = z < 1 ? 1 : (z > 256 ? 256 : z);
  *qround = r < 1 ? 1 : (r > 256 ? 256 : r);
}

/* enc/quant.rs:195 build_quant_table_bd_sharp */
void or_build_quant_row_sharp(int qindex, int bd, int sharpness, or_quant_row *t) {
  int32_t dc = bd == 10 ? DC_QLOOKUP_10[qindex] : DC_QLOOKUP_8[qindex];
  int32_t ac = bd == 10 ? AC_QLOOKUP_10[qindex] : AC_QLOOKUP_8[qindex];
  int th = bd == 10 ? 592 : 148;
  int qzbin = qindex == 0 ? 64 : (dc < th ? 84 : 80);
  int qround = qindex == 0 ? 64 : 48;
  sharp_factors(&qzbin, &qround, qindex, sharpness);
  int32_t q[2] = {dc, ac};
  for (int i = 0; i < 2; i++) {
    invert_quant(q[i], &t->quant[i], &t->quant_shift[i]);
    t->zbin[i] = (qzbin * q[i] + 64) >> 7;
    t->round[i] = (qround * q[i]) >> 7;
    t->quant_fp[i] = (1 << 16) / q[i];
    t->round_fp[i] = (64 * q[i]) >> 7;
    t->dequant[i] = q[i];
  }
}

void or_build_quant_row(int qindex, int bd, or_quant_row *t) {
  int32_t dc = bd == 10 ? DC_QLOOKUP_10[qindex] : DC_QLOOKUP_8[qindex];
  int32_t ac = bd == 10 ? AC_QLOOKUP_10[qindex] : AC_QLOOKUP_8[qindex];
  int th = bd == 10 ? 592 : 148;
  int qzbin = qindex == 0 ? 64 : (dc < th ? 84 : 80);
  int qround = qindex == 0 ? 64 : 48;
  int32_t q[2] = {dc, ac};
  for (int i = 0; i < 2; i++) {
    invert_quant(q[i], &t->quant[i], &t->quant_shift[i]);
    t->zbin[i] = (qzbin * q[i] + 64) >> 7;
    t->round[i] = (qround * q[i]) >> 7;
    t->quant_fp[i] = (1 << 16) / q[i];
    t->round_fp[i] = (64 * q[i]) >> 7;
    t->dequant[i] = q[i];
  }
}

static int64_t clamp64(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* dsp/quant_coding.rs:80 fp_one; enc/quant.rs:286 (bd8, INT16 clamp) and
 * :335 quantize_fp_hbd (bd10, no clamp).  Raster order; position 0 is DC. */
void or_quantize_fp(const int32_t *coeff, int n, const or_quant_row *t, int ls, int bd,
                    int32_t *qcoeff, int32_t *dqcoeff) {
  int32_t rounding[2];
  for (int i = 0; i < 2; i++) rounding[i] = (t->round_fp[i] + ((1 << ls) >> 1)) >> ls;
  for (int rc = 0; rc < n; rc++) {
    int iz = rc != 0;
    int32_t c = coeff[rc];
    int32_t sgn = c < 0 ? -1 : 0;
    int64_t a = (int64_t)((c ^ sgn) - sgn);
    int32_t tmp32 = 0;
    if ((a << (1 + ls)) >= (int64_t)t->dequant[iz]) {
      int64_t x = a + rounding[iz];
      if (bd == 8) x = clamp64(x, -32768, 32767);
      tmp32 = (int32_t)((x * t->quant_fp[iz]) >> (16 - ls));
    }
    if (tmp32) {
      qcoeff[rc] = (tmp32 ^ sgn) - sgn;
      int32_t adq = (int32_t)(((int64_t)tmp32 * t->dequant[iz]) >> ls);
      dqcoeff[rc] = (adq ^ sgn) - sgn;
    } else {
      qcoeff[rc] = 0;
      dqcoeff[rc] = 0;
    }
  }
}

/* dsp/quant_coding.rs:365 b_one; enc/quant.rs:247 (bd8 clamp) and :391
 * quantize_b_hbd (bd10, no clamp; its prescan is outcome-identical). */
void or_quantize_b(const int32_t *coeff, int n, const or_quant_row *t, int ls, int bd,
                   int32_t *qcoeff, int32_t *dqcoeff) {
  int32_t zbin[2], round[2];
  for (int i = 0; i < 2; i++) {
    zbin[i] = (t->zbin[i] + ((1 << ls) >> 1)) >> ls;
    round[i] = (t->round[i] + ((1 << ls) >> 1)) >> ls;
  }
  for (int rc = 0; rc < n; rc++) {
    int iz = rc != 0;
    int32_t c = coeff[rc];
    int32_t sgn = c < 0 ? -1 : 0;
    int32_t a = (c ^ sgn) - sgn;
    int32_t tmp32 = 0;
    if (a >= zbin[iz]) {
      int64_t x = (int64_t)a + round[iz];
      if (bd == 8) x = clamp64(x, -32768, 32767);
      tmp32 = (int32_t)(((((x * t->quant[iz]) >> 16) + x) * t->quant_shift[iz]) >> (16 - ls));
    }
    if (tmp32) {
      qcoeff[rc] = (tmp32 ^ sgn) - sgn;
      int32_t adq = (int32_t)(((int64_t)tmp32 * t->dequant[iz]) >> ls);
      dqcoeff[rc] = (adq ^ sgn) - sgn;
    } else {
      qcoeff[rc] = 0;
      dqcoeff[rc] = 0;
    }
  }
}

/* dsp/residual.rs:250-283 sse_i32_core: i64 difference, wrapping square and sum. */
uint64_t or_sse_i32(const int32_t *a, const int32_t *b, int n) {
  uint64_t d = 0;
  for (int i = 0; i < n; i++) {
    int64_t e = (int64_t)a[i] - (int64_t)b[i];
    d += (uint64_t)e * (uint64_t)e;
  }
  return d;
}

/* dsp/residual.rs:374 sq_sum_i32_core */
uint64_t or_sq_sum_i32(const int32_t *a, int n) {
  uint64_t d = 0;
  for (int i = 0; i < n; i++) d += (uint64_t)((int64_t)a[i] * a[i]);
  return d;
}

/* dsp/variance.rs:124 sse_impl_scalar; dsp/hbd.rs:936 for u16 samples. */
uint64_t or_sse_pix(const void *src, int ss, const void *ref, int rs, int w, int h, int hbd) {
  uint64_t d = 0;
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      int32_t a = hbd ? ((const uint16_t *)src)[r * ss + c] : ((const uint8_t *)src)[r * ss + c];
      int32_t b = hbd ? ((const uint16_t *)ref)[r * rs + c] : ((const uint8_t *)ref)[r * rs + c];
      int32_t e = a - b;
      d += (uint64_t)(int64_t)(e * e);
    }
  return d;
}

/* dsp/hadamard.rs:643 hadamard_col8 (C output permutation, i16 store). */
static void had_col8(const int16_t *s, int st, int16_t *o) {
  int32_t b0 = s[0] + s[st], b1 = s[0] - s[st];
  int32_t b2 = s[2 * st] + s[3 * st], b3 = s[2 * st] - s[3 * st];
  int32_t b4 = s[4 * st] + s[5 * st], b5 = s[4 * st] - s[5 * st];
  int32_t b6 = s[6 * st] + s[7 * st], b7 = s[6 * st] - s[7 * st];
  int32_t c0 = b0 + b2, c1 = b1 + b3, c2 = b0 - b2, c3 = b1 - b3;
  int32_t c4 = b4 + b6, c5 = b5 + b7, c6 = b4 - b6, c7 = b5 - b7;
  o[0] = (int16_t)(c0 + c4); o[7] = (int16_t)(c1 + c5);
  o[3] = (int16_t)(c2 + c6); o[4] = (int16_t)(c3 + c7);
  o[2] = (int16_t)(c0 - c4); o[6] = (int16_t)(c1 - c5);
  o[1] = (int16_t)(c2 - c6); o[5] = (int16_t)(c3 - c7);
}

/* dsp/hadamard.rs:675 hadamard_col4 (i16 arithmetic with >>1 first stage). */
static void had_col4(const int16_t *s, int st, int16_t *o) {
  int16_t b0 = (int16_t)((int16_t)(s[0] + s[st]) >> 1);
  int16_t b1 = (int16_t)((int16_t)(s[0] - s[st]) >> 1);
  int16_t b2 = (int16_t)((int16_t)(s[2 * st] + s[3 * st]) >> 1);
  int16_t b3 = (int16_t)((int16_t)(s[2 * st] - s[3 * st]) >> 1);
  o[0] = (int16_t)(b0 + b2); o[1] = (int16_t)(b1 + b3);
  o[2] = (int16_t)(b0 - b2); o[3] = (int16_t)(b1 - b3);
}

static void had4(const int16_t *src, int stride, int32_t *coeff) {
  int16_t b1[16], b2[16];
  for (int i = 0; i < 4; i++) had_col4(src + i, stride, b1 + 4 * i);
  for (int i = 0; i < 4; i++) had_col4(b1 + i, 4, b2 + 4 * i);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) coeff[i * 4 + j] = b2[j * 4 + i];
}

static void had8(const int16_t *src, int stride, int32_t *coeff) {
  int16_t b1[64], b2[64];
  for (int i = 0; i < 8; i++) had_col8(src + i, stride, b1 + 8 * i);
  for (int i = 0; i < 8; i++) had_col8(b1 + i, 8, b2 + 8 * i);
  for (int i = 0; i < 64; i++) coeff[i] = b2[i];
}

/* dsp/hadamard.rs:885 aom_hadamard_16x16 (AVX2 16-bit wrap semantics). */
static void had16(const int16_t *src, int stride, int32_t *coeff) {
  for (int q = 0; q < 4; q++) had8(src + (q >> 1) * 8 * stride + (q & 1) * 8, stride, coeff + q * 64);
  for (int i = 0; i < 64; i++) {
    int16_t a0 = (int16_t)coeff[i], a1 = (int16_t)coeff[i + 64];
    int16_t a2 = (int16_t)coeff[i + 128], a3 = (int16_t)coeff[i + 192];
    int16_t b0 = (int16_t)((int16_t)(a0 + a1) >> 1), b1 = (int16_t)((int16_t)(a0 - a1) >> 1);
    int16_t b2 = (int16_t)((int16_t)(a2 + a3) >> 1), b3 = (int16_t)((int16_t)(a2 - a3) >> 1);
    coeff[i] = (int16_t)(b0 + b2);
    coeff[i + 64] = (int16_t)(b1 + b3);
    coeff[i + 128] = (int16_t)(b0 - b2);
    coeff[i + 192] = (int16_t)(b1 - b3);
  }
}

static int16_t sat16(int32_t v) { return (int16_t)(v < -32768 ? -32768 : (v > 32767 ? 32767 : v)); }

/* dsp/hadamard.rs:913 aom_hadamard_32x32 (>>2 in 32-bit, saturate, wrap). */
static void had32(const int16_t *src, int stride, int32_t *coeff) {
  for (int q = 0; q < 4; q++) had16(src + (q >> 1) * 16 * stride + (q & 1) * 16, stride, coeff + q * 256);
  for (int i = 0; i < 256; i++) {
    int32_t a0 = (int16_t)coeff[i], a1 = (int16_t)coeff[i + 256];
    int32_t a2 = (int16_t)coeff[i + 512], a3 = (int16_t)coeff[i + 768];
    int16_t b0 = sat16((a0 + a1) >> 2), b1 = sat16((a0 - a1) >> 2);
    int16_t b2 = sat16((a2 + a3) >> 2), b3 = sat16((a2 - a3) >> 2);
    coeff[i] = (int16_t)(b0 + b2);
    coeff[i + 256] = (int16_t)(b1 + b3);
    coeff[i + 512] = (int16_t)(b0 - b2);
    coeff[i + 768] = (int16_t)(b1 - b3);
  }
}

void or_hadamard(int n, const int16_t *src, int stride, int32_t *coeff) {
  switch (n) {
    case 4: had4(src, stride, coeff); break;
    case 8: had8(src, stride, coeff); break;
    case 16: had16(src, stride, coeff); break;
    case 32: had32(src, stride, coeff); break;
    default: break;
  }
}

/* dsp/hadamard.rs:942 aom_satd */
int32_t or_satd(const int32_t *coeff, int n) {
  int32_t s = 0;
  for (int i = 0; i < n; i++) s += coeff[i] < 0 ? -coeff[i] : coeff[i];
  return s;
}

/* lf/predict.rs:139 hadamard_satd / :185 _hbd.  The residual is formed in
 * i16 (u8/u16 sample cast to i16, then subtracted). */
uint64_t or_block_satd(const void *src, int ss, const void *pred, int ps, int w, int h, int hbd) {
  int tx = w < h ? w : h;
  if (tx > 32) tx = 32;
  int16_t res[32 * 32];
  int32_t coeff[32 * 32];
  uint64_t satd = 0;
  for (int ty = 0; ty < h; ty += tx)
    for (int tx0 = 0; tx0 < w; tx0 += tx) {
      for (int r = 0; r < tx; r++)
        for (int c = 0; c < tx; c++) {
          int off_s = (ty + r) * ss + tx0 + c, off_p = (ty + r) * ps + tx0 + c;
          int16_t a = hbd ? (int16_t)((const uint16_t *)src)[off_s] : (int16_t)((const uint8_t *)src)[off_s];
          int16_t b = hbd ? (int16_t)((const uint16_t *)pred)[off_p] : (int16_t)((const uint8_t *)pred)[off_p];
          res[r * tx + c] = (int16_t)(a - b);
        }
      or_hadamard(tx, res, tx, coeff);
      satd += (uint64_t)(uint32_t)or_satd(coeff, tx * tx);
    }
  return satd;
}
