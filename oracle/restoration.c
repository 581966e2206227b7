/*
 * Oracle: loop-restoration statistics (TEST INFRASTRUCTURE ONLY — see
 * zav1_oracle.h).
 *
 *   find_average(_hbd)        dsp/restoration.rs:198 / :1707
 *   compute_stats(_hbd)       dsp/restoration.rs:232 + scalar core :567 (bd8) / :1738 (hbd: i64
 *                             products, / 4 at bd10 after accumulation)
 *   unit iteration            dsp/restoration.rs:1388 count_units_in_tile,
 *                             :1395 foreach_rest_unit_in_tile
 *   dgd border                enc/restoration.rs:740 extend_frame (edge
 *                             replication)
 *   sse_region(_hbd)          dsp/restoration.rs:1666 / :1991
 *
 * Self-guided statistics have NO counterpart in the reference
 * (enc/restoration.rs:5-9); the restatement below follows the algorithm
 * text of specs/08-loop-filters.md:371-420 and the AV1 specification's
 * self-guided filter process.  No reference artifact pins it: it is checked
 * against the specification's parameter sets and lookup tables and a
 * whole-array numpy brute force of every step
 * (tests/test_oracle_pin_lr_cdef.py), and GPU == this file.
 */
#include <stdlib.h>
#include <string.h>

#include "zav1_oracle.h"

static int32_t S(const void *p, int hbd, long i) { return hbd ? ((const uint16_t *)p)[i] : ((const uint8_t *)p)[i]; }

/* Replicated-border sample of a tight plane (extend_frame semantics). */
static int32_t E(const void *p, int hbd, int pw, int ph, int x, int y) {
  x = x < 0 ? 0 : (x >= pw ? pw - 1 : x);
  y = y < 0 ? 0 : (y >= ph ? ph - 1 : y);
  return S(p, hbd, (long)y * pw + x);
}

int or_count_units(int unit_size, int size) {
  int n = (size + (unit_size >> 1)) / unit_size;
  return n < 1 ? 1 : n;
}

/* foreach_rest_unit_in_tile for a whole plane: limits of unit (ux, uy). */
void or_unit_limits(int unit_size, int pw, int ph, int ss_y, int ux, int uy, int *hs, int *he, int *vs, int *ve) {
  int ext = unit_size * 3 / 2;
  int y0 = uy * unit_size, x0 = ux * unit_size;
  int h = (ph - y0 < ext) ? ph - y0 : unit_size;
  int w = (pw - x0 < ext) ? pw - x0 : unit_size;
  int voff = 8 >> ss_y;
  *vs = y0 - voff < 0 ? 0 : y0 - voff;
  *ve = y0 + h;
  if (*ve < ph) *ve -= voff;
  *hs = x0;
  *he = x0 + w;
}

/* compute_stats / compute_stats_hbd over [hs,he) x [vs,ve); dgd and src are
 * tight pw x ph planes, dgd read with border replication. */
void or_wiener_stats(int win, const void *dgd, const void *src, int pw, int ph, int bd, int hs, int he, int vs,
                     int ve, int64_t *M, int64_t *H) {
  const int hbd = bd > 8, w2 = win * win, half = win >> 1;
  uint64_t sum = 0;
  for (int i = vs; i < ve; i++)
    for (int j = hs; j < he; j++) sum += (uint64_t)E(dgd, hbd, pw, ph, j, i);
  const int32_t avg = (int32_t)(sum / ((uint64_t)(ve - vs) * (uint64_t)(he - hs)));
  memset(M, 0, sizeof(int64_t) * w2);
  memset(H, 0, sizeof(int64_t) * w2 * w2);
  int32_t y[49];
  for (int i = vs; i < ve; i++)
    for (int j = hs; j < he; j++) {
      const int64_t x = S(src, hbd, (long)i * pw + j) - avg;
      int idx = 0;
      for (int k = -half; k <= half; k++)
        for (int l = -half; l <= half; l++) y[idx++] = E(dgd, hbd, pw, ph, j + k, i + l) - avg;
      for (int k = 0; k < w2; k++) {
        M[k] += (int64_t)y[k] * x;
        for (int l = k; l < w2; l++) H[k * w2 + l] += (int64_t)y[k] * y[l];
      }
    }
  const int64_t div = bd == 10 ? 4 : 1;
  for (int k = 0; k < w2; k++) {
    M[k] /= div;
    H[k * w2 + k] /= div;
    for (int l = k + 1; l < w2; l++) {
      H[k * w2 + l] /= div;
      H[l * w2 + k] = H[k * w2 + l];
    }
  }
}

/* sse_region(_hbd): SSE of the unfiltered reconstruction over the unit. */
uint64_t or_sse_region(const void *a, const void *b, int pw, int bd, int hs, int he, int vs, int ve) {
  const int hbd = bd > 8;
  uint64_t d = 0;
  for (int i = vs; i < ve; i++)
    for (int j = hs; j < he; j++) {
      int64_t e = (int64_t)S(a, hbd, (long)i * pw + j) - S(b, hbd, (long)i * pw + j);
      d += (uint64_t)(e * e);
    }
  return d;
}

/* ------------------------------------------------------------------ */
/* self-guided filter + projection statistics (spec-derived)           */
/* ------------------------------------------------------------------ */
/* AV1 Sgr_Params: {r0, s0, r1, s1} for the 16 parameter sets. */
static const int SGR_PARAMS[16][4] = {
    {2, 140, 1, 3236}, {2, 112, 1, 2158}, {2, 93, 1, 1618}, {2, 80, 1, 1438}, {2, 70, 1, 1295}, {2, 58, 1, 1177},
    {2, 47, 1, 1079},  {2, 37, 1, 996},   {2, 30, 1, 925},  {2, 25, 1, 863},  {0, -1, 1, 2589}, {0, -1, 1, 1618},
    {0, -1, 1, 1177},  {0, -1, 1, 925},   {2, 56, 0, -1},   {2, 22, 0, -1}};

static int x_by_xplus1(int z) { return z == 0 ? 1 : (z >= 255 ? 256 : (256 * z + (z + 1) / 2) / (z + 1)); }
static int one_by_x(int n) { return (4096 + n / 2) / n; }
static int32_t rpot(int64_t v, int n) { return n == 0 ? (int32_t)v : (int32_t)((v + ((int64_t)1 << (n - 1))) >> n); }

int or_sgr_param(int set, int k) { return SGR_PARAMS[set][k]; }
/* the two lookup tables of the self-guided filter: 0 = x_by_xplus1[z]
 * (z < 256), 1 = one_by_x[n - 1] (n = 1..25) */
int or_sgr_table(int which, int i) { return which == 0 ? x_by_xplus1(i) : one_by_x(i + 1); }

/* One self-guided pass of radius r (1 or 2) and scale s over the unit
 * [hs,he) x [vs,ve) of the replicated-border plane; flt is (he-hs) x (ve-vs)
 * in units of 1 << 4 (SGRPROJ_RST_BITS). */
static void sgr_pass(const void *dgd, int hbd, int pw, int ph, int bd, int hs, int he, int vs, int ve, int r, int s,
                     int32_t *flt) {
  const int W = he - hs, Hh = ve - vs, n = (2 * r + 1) * (2 * r + 1);
  /* A/B over rows/cols -1 .. H/W (buffer offset by 1) */
  const int bw = W + 2, bh = Hh + 2;
  int32_t *A = (int32_t *)calloc((size_t)bw * bh, sizeof(int32_t));
  int32_t *B = (int32_t *)calloc((size_t)bw * bh, sizeof(int32_t));
  for (int i = -1; i < Hh + 1; i++) {
    if (r == 2 && ((i + 1) & 1)) continue; /* fast variant: rows -1, 1, 3, ... */
    for (int j = -1; j < W + 1; j++) {
      int64_t sum = 0, sq = 0;
      for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
          int32_t v = E(dgd, hbd, pw, ph, hs + j + dx, vs + i + dy);
          sum += v;
          sq += (int64_t)v * v;
        }
      int32_t a = rpot(sq, 2 * (bd - 8)), b = rpot(sum, bd - 8);
      int64_t p = (int64_t)a * n - (int64_t)b * b;
      if (p < 0) p = 0;
      uint32_t z = (uint32_t)rpot(p * s, 20);
      int32_t av = x_by_xplus1(z > 255 ? 255 : (int)z);
      int32_t bv = rpot((int64_t)(uint32_t)(256 - av) * (uint32_t)sum * (uint32_t)one_by_x(n), 12);
      A[(i + 1) * bw + j + 1] = av;
      B[(i + 1) * bw + j + 1] = bv;
    }
  }
#define AT(X, yy, xx) X[((yy) + 1) * bw + (xx) + 1]
  for (int i = 0; i < Hh; i++)
    for (int j = 0; j < W; j++) {
      int32_t d = E(dgd, hbd, pw, ph, hs + j, vs + i);
      int64_t a, b;
      int shift;
      if (r == 1) {
        a = (int64_t)(AT(A, i, j) + AT(A, i, j - 1) + AT(A, i, j + 1) + AT(A, i - 1, j) + AT(A, i + 1, j)) * 4 +
            (int64_t)(AT(A, i - 1, j - 1) + AT(A, i - 1, j + 1) + AT(A, i + 1, j - 1) + AT(A, i + 1, j + 1)) * 3;
        b = (int64_t)(AT(B, i, j) + AT(B, i, j - 1) + AT(B, i, j + 1) + AT(B, i - 1, j) + AT(B, i + 1, j)) * 4 +
            (int64_t)(AT(B, i - 1, j - 1) + AT(B, i - 1, j + 1) + AT(B, i + 1, j - 1) + AT(B, i + 1, j + 1)) * 3;
        shift = 8 + 5 - 4;
      } else if ((i & 1) == 0) {
        a = (int64_t)(AT(A, i - 1, j) + AT(A, i + 1, j)) * 6 +
            (int64_t)(AT(A, i - 1, j - 1) + AT(A, i - 1, j + 1) + AT(A, i + 1, j - 1) + AT(A, i + 1, j + 1)) * 5;
        b = (int64_t)(AT(B, i - 1, j) + AT(B, i + 1, j)) * 6 +
            (int64_t)(AT(B, i - 1, j - 1) + AT(B, i - 1, j + 1) + AT(B, i + 1, j - 1) + AT(B, i + 1, j + 1)) * 5;
        shift = 8 + 5 - 4;
      } else {
        a = (int64_t)AT(A, i, j) * 6 + (int64_t)(AT(A, i, j - 1) + AT(A, i, j + 1)) * 5;
        b = (int64_t)AT(B, i, j) * 6 + (int64_t)(AT(B, i, j - 1) + AT(B, i, j + 1)) * 5;
        shift = 8 + 4 - 4;
      }
      flt[i * W + j] = rpot(a * d + b, shift);
    }
#undef AT
  free(A);
  free(B);
}

/* Projection normal equations of parameter set `set` for one unit:
 * out[0..5) = H00, H01, H11, C0, C1 (sums divided by the unit area, as the
 * projection solver consumes them). */
void or_sgr_unit_stats(const void *dgd, const void *src, int pw, int ph, int bd, int hs, int he, int vs, int ve,
                       int set, int64_t *out) {
  const int hbd = bd > 8, W = he - hs, Hh = ve - vs, size = W * Hh;
  const int r0 = SGR_PARAMS[set][0], s0 = SGR_PARAMS[set][1], r1 = SGR_PARAMS[set][2], s1 = SGR_PARAMS[set][3];
  int32_t *f0 = (int32_t *)calloc((size_t)size, sizeof(int32_t));
  int32_t *f1 = (int32_t *)calloc((size_t)size, sizeof(int32_t));
  if (r0) sgr_pass(dgd, hbd, pw, ph, bd, hs, he, vs, ve, r0, s0, f0);
  if (r1) sgr_pass(dgd, hbd, pw, ph, bd, hs, he, vs, ve, r1, s1, f1);
  int64_t h00 = 0, h01 = 0, h11 = 0, c0 = 0, c1 = 0;
  for (int i = 0; i < Hh; i++)
    for (int j = 0; j < W; j++) {
      const int64_t u = (int64_t)S(dgd, hbd, (long)(vs + i) * pw + hs + j) << 4;
      const int64_t sv = ((int64_t)S(src, hbd, (long)(vs + i) * pw + hs + j) << 4) - u;
      const int64_t a = r0 ? f0[i * W + j] - u : 0, b = r1 ? f1[i * W + j] - u : 0;
      h00 += a * a;
      h01 += a * b;
      h11 += b * b;
      c0 += a * sv;
      c1 += b * sv;
    }
  out[0] = h00 / size;
  out[1] = h01 / size;
  out[2] = h11 / size;
  out[3] = c0 / size;
  out[4] = c1 / size;
  free(f0);
  free(f1);
}
