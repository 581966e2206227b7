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
 *   Wiener unit filter        dsp/restoration.rs:138 / :1820 wiener_convolve_add_src(_hbd),
 *                             :1288 loop_restoration_filter_unit (search path),
 *                             :1921 loop_restoration_filter_unit_search_hbd
 *   finer tap search          enc/restoration.rs:513 finer_tile_search_wiener
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
static uint64_t sse_rect(const void *a, const void *b, int hbd, int pw, int x0, int x1, int y0, int y1) {
  uint64_t d = 0;
  for (int i = y0; i < y1; i++)
    for (int j = x0; j < x1; j++) {
      int64_t e = (int64_t)S(a, hbd, (long)i * pw + j) - S(b, hbd, (long)i * pw + j);
      d += (uint64_t)(e * e);
    }
  return d;
}

/* sse_region: the plain sum at 8 bits; at 10 bits highbd_variance's parts,
 * each truncated to 32 bits before the sum (dsp/restoration.rs:1991): the
 * right strip of width % 16 columns over every row, the bottom strip of
 * height % 16 rows left of it, and the 16x16 blocks. */
uint64_t or_sse_region(const void *a, const void *b, int pw, int bd, int hs, int he, int vs, int ve) {
  const int hbd = bd > 8;
  if (!hbd) return sse_rect(a, b, 0, pw, hs, he, vs, ve);
  const int w = he - hs, h = ve - vs, dw = w % 16, dh = h % 16;
  uint64_t t = 0;
  if (dw) t += (uint32_t)sse_rect(a, b, 1, pw, he - dw, he, vs, ve);
  if (dh) t += (uint32_t)sse_rect(a, b, 1, pw, hs, he - dw, ve - dh, ve);
  for (int y = 0; y < h / 16; y++)
    for (int x = 0; x < w / 16; x++)
      t += (uint32_t)sse_rect(a, b, 1, pw, hs + 16 * x, hs + 16 * x + 16, vs + 16 * y, vs + 16 * y + 16);
  return t;
}

/* ------------------------------------------------------------------ */
/* Wiener unit filter and the finer tap search                          */
/* ------------------------------------------------------------------ */
/* wiener_convolve_add_src(_hbd) (dsp/restoration.rs:138 / :1820) at one
 * output pixel of the edge-extended plane: the horizontal pass over rows
 * y - 3 .. y + 3 (round 3, clamped to [0, 2^(bd + 5) - 1]), then the
 * vertical pass (round 11, clamped to the pixel range).  The 8th taps are
 * 0.  Per output pixel this is what the stripe / processing-unit loops of
 * loop_restoration_filter_unit (:1288, need_boundaries = 0) and
 * loop_restoration_filter_unit_search_hbd (:1921) compute. */
static int wiener_px(const void *dgd, int hbd, int bd, int pw, int ph, int x, int y, const int16_t *hf,
                     const int16_t *vf) {
  const int clamp_h = (1 << (bd + 1 + 7 - 3)) - 1, mx = (1 << bd) - 1;
  int32_t t[7];
  for (int r = 0; r < 7; r++) {
    int32_t sum = (E(dgd, hbd, pw, ph, x, y + r - 3) << 7) + (1 << (bd + 7 - 1));
    for (int k = 0; k < 7; k++) sum += E(dgd, hbd, pw, ph, x + k - 3, y + r - 3) * hf[k];
    int32_t v = (sum + 4) >> 3;
    t[r] = v < 0 ? 0 : (v > clamp_h ? clamp_h : v);
  }
  int32_t sum = (t[3] << 7) - (1 << (bd + 11 - 1));
  for (int k = 0; k < 7; k++) sum += t[k] * vf[k];
  int32_t v = (sum + 1024) >> 11;
  return v < 0 ? 0 : (v > mx ? mx : v);
}

/* The unit filter over every restoration unit of a plane
 * (loop_restoration_filter_unit without stripe boundaries, as the search
 * runs it): types[u] 0 copies the unit, 1 filters it with filters[u * 16]
 * (hfilter[8]) and [u * 16 + 8] (vfilter[8]). */
void or_wiener_filter_plane(const void *dgd, void *out, int pw, int ph, int bd, int unit_size, int ss_y,
                            const uint8_t *types, const int16_t *filters) {
  const int hbd = bd > 8, nh = or_count_units(unit_size, pw), nv = or_count_units(unit_size, ph);
  for (int uy = 0; uy < nv; uy++)
    for (int ux = 0; ux < nh; ux++) {
      int hs, he, vs, ve;
      or_unit_limits(unit_size, pw, ph, ss_y, ux, uy, &hs, &he, &vs, &ve);
      const int u = uy * nh + ux;
      for (int y = vs; y < ve; y++)
        for (int x = hs; x < he; x++) {
          const int v = types[u] ? wiener_px(dgd, hbd, bd, pw, ph, x, y, filters + u * 16, filters + u * 16 + 8)
                                 : S(dgd, hbd, (long)y * pw + x);
          if (hbd) ((uint16_t *)out)[(long)y * pw + x] = (uint16_t)v;
          else ((uint8_t *)out)[(long)y * pw + x] = (uint8_t)v;
        }
    }
}

/* try_restoration_unit (enc/restoration.rs:472): the unit filtered with
 * (hf, vf), SSE against the source over the unit. */
static int64_t try_unit(const void *dgd, const void *src, int hbd, int bd, int pw, int ph, int hs, int he, int vs,
                        int ve, const int16_t *hf, const int16_t *vf, void *trial) {
  for (int y = vs; y < ve; y++)
    for (int x = hs; x < he; x++) {
      const int v = wiener_px(dgd, hbd, bd, pw, ph, x, y, hf, vf);
      if (hbd) ((uint16_t *)trial)[(long)y * pw + x] = (uint16_t)v;
      else ((uint8_t *)trial)[(long)y * pw + x] = (uint8_t)v;
    }
  return (int64_t)or_sse_region(src, trial, pw, bd, hs, he, vs, ve);
}

/* finer_tile_search_wiener (enc/restoration.rs:513): the unit's SSE with
 * (hf, vf), then (use_refinement) the +-step hill climb over the taps
 * plane_off .. 2 of hfilter, then of vfilter, steps 4, 2, 1 (4 only with
 * max_one_step), each tap moved with its mirror and the centre rebalanced,
 * kept when the SSE does not grow; at the start step a kept move is
 * repeated, and a kept minus move ends that pass over the taps (the
 * reference's break out of the tap loop).  Returns the final SSE; hf / vf
 * are updated in place. */
int64_t or_wiener_refine_unit(const void *dgd, const void *src, int pw, int ph, int bd, int win, int hs, int he,
                              int vs, int ve, int use_refinement, int max_one_step, int16_t *hf, int16_t *vf) {
  static const int16_t tmin[3] = {3 - 8, -7 - 16, 15 - 32}, tmax[3] = {3 - 1 + 8, -7 - 1 + 16, 15 - 1 + 32};
  const int hbd = bd > 8;
  void *trial = calloc((size_t)pw * ph, hbd ? 2 : 1);
  int64_t err = try_unit(dgd, src, hbd, bd, pw, ph, hs, he, vs, ve, hf, vf, trial);
  if (use_refinement) {
    const int plane_off = (7 - win) >> 1, start = 4, end = max_one_step ? 4 : 1;
    for (int s = start; s >= end; s >>= 1)
      for (int pass = 0; pass < 2; pass++) {
        int16_t *f = pass ? vf : hf;
        for (int p = plane_off; p < 3; p++) {
          int skip = 0;
          for (;;) { /* minus */
            if (f[p] - s >= tmin[p]) {
              f[p] -= s, f[6 - p] -= s, f[3] += 2 * s;
              const int64_t e2 = try_unit(dgd, src, hbd, bd, pw, ph, hs, he, vs, ve, hf, vf, trial);
              if (e2 > err) {
                f[p] += s, f[6 - p] += s, f[3] -= 2 * s;
              } else {
                err = e2;
                skip = 1;
                if (s == start && !max_one_step) continue;
              }
            }
            break;
          }
          if (skip) break; /* a kept minus move ends the pass over the taps */
          for (;;) { /* plus */
            if (f[p] + s <= tmax[p]) {
              f[p] += s, f[6 - p] += s, f[3] -= 2 * s;
              const int64_t e2 = try_unit(dgd, src, hbd, bd, pw, ph, hs, he, vs, ve, hf, vf, trial);
              if (e2 > err) {
                f[p] -= s, f[6 - p] -= s, f[3] += 2 * s;
              } else {
                err = e2;
                if (s == start && !max_one_step) continue;
              }
            }
            break;
          }
        }
      }
  }
  free(trial);
  return err;
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
