/*
 * Oracle: deblocking filters + frame edge walk, and the CDEF direction /
 * filter / strength-SSE search (TEST INFRASTRUCTURE ONLY — see
 * zav1_oracle.h).
 *
 *   lf_thresholds                dsp/loop_filter.rs:50
 *   masks, filter4/6/8/14        dsp/loop_filter.rs:76-286; hbd twins
 *                                dsp/hbd.rs:1040-1330 (8-bit is the bd=8 case:
 *                                bias 0x80, clamp [-128,127], no shift)
 *   lpf_params / filter_plane    enc/deblock.rs:698 / :811 (all vertical
 *                                edges, then all horizontal edges)
 *   cdef_find_dir                dsp/cdef.rs:143
 *   constrain / adjust_strength  dsp/cdef.rs:113 / :125
 *   cdef_filter_block            dsp/cdef.rs:931 (hbd dsp/hbd.rs:1765)
 *   build_src / filter_fb_packed / dist / cdef_search_still(_hbd)
 *                                enc/cdef.rs:517 / :952,:1045 / :1128 /
 *                                :1166,:1374
 */
#include <stdlib.h>
#include <string.h>

#include "zav1_oracle.h"

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

/* ------------------------------------------------------------------ */
/* deblocking                                                          */
/* ------------------------------------------------------------------ */
void or_lf_thresholds(int level, int sharpness, int *mblim, int *lim, int *hev) {
  int bil = level >> ((sharpness > 0) + (sharpness > 4));
  if (sharpness > 0 && bil > 9 - sharpness) bil = 9 - sharpness;
  if (bil < 1) bil = 1;
  *mblim = 2 * (level + 2) + bil;
  *lim = bil;
  *hev = level >> 4;
}

typedef struct { int32_t lim, mblim, hev, shift, bd; } thr_t;

static int32_t sclamp(int32_t t, int bd) {
  int32_t lo = -(128 << (bd - 8)), hi = (128 << (bd - 8)) - 1;
  return clampi(t, lo, hi);
}

/* filter_mask over n samples w[0..n) (n = 4, 6, 8): limit on neighbour
 * steps, blimit on the p0/q0 step. */
static int mask_n(const thr_t *t, const int32_t *w, int n) {
  int32_t l = t->lim << t->shift, bl = t->mblim << t->shift;
  int h = n / 2, bad = 0;
  for (int k = 0; k + 1 < h; k++) bad |= iabs(w[k] - w[k + 1]) > l;          /* p side */
  for (int k = h; k + 1 < n; k++) bad |= iabs(w[k] - w[k + 1]) > l;          /* q side */
  bad |= iabs(w[h - 1] - w[h]) * 2 + iabs(w[h - 2] - w[h + 1]) / 2 > bl;
  return !bad;
}

static int flat_n(int32_t thr, int shift, const int32_t *p, const int32_t *q, int m) {
  /* |p_k - p0| and |q_k - q0| for k = 1..m-1 against thr << shift */
  int32_t t = thr << shift;
  int bad = 0;
  for (int k = 1; k < m; k++) bad |= iabs(p[k] - p[0]) > t || iabs(q[k] - q[0]) > t;
  return !bad;
}

static void filter4(int mask, const thr_t *t, int32_t *w) {
  int32_t bias = 0x80 << t->shift;
  int32_t ps1 = w[0] - bias, ps0 = w[1] - bias, qs0 = w[2] - bias, qs1 = w[3] - bias;
  int32_t th = t->hev << t->shift;
  int32_t hev = (iabs(w[0] - w[1]) > th || iabs(w[3] - w[2]) > th) ? -1 : 0;
  int32_t m = mask ? -1 : 0;
  int32_t f = sclamp(ps1 - qs1, t->bd) & hev;
  f = sclamp(f + 3 * (qs0 - ps0), t->bd) & m;
  int32_t f1 = sclamp(f + 4, t->bd) >> 3, f2 = sclamp(f + 3, t->bd) >> 3;
  w[2] = sclamp(qs0 - f1, t->bd) + bias;
  w[1] = sclamp(ps0 + f2, t->bd) + bias;
  int32_t o = ((f1 + 1) >> 1) & ~hev;
  w[3] = sclamp(qs1 - o, t->bd) + bias;
  w[0] = sclamp(ps1 + o, t->bd) + bias;
}

#define R3(x) (((x) + 4) >> 3)
#define R4(x) (((x) + 8) >> 4)

/* One line across an edge: w holds `len` samples centred on the edge. */
static void lpf_line(int len, const thr_t *t, int32_t *w) {
  if (len == 4) {
    filter4(mask_n(t, w, 4), t, w);
  } else if (len == 6) {
    int m = mask_n(t, w, 6);
    int32_t p[3] = {w[2], w[1], w[0]}, q[3] = {w[3], w[4], w[5]};
    int fl = flat_n(1, t->shift, p, q, 3);
    if (fl && m) {
      int32_t p2 = w[0], p1 = w[1], p0 = w[2], q0 = w[3], q1 = w[4], q2 = w[5];
      w[1] = R3(p2 * 3 + p1 * 2 + p0 * 2 + q0);
      w[2] = R3(p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1);
      w[3] = R3(p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2);
      w[4] = R3(p0 + q0 * 2 + q1 * 2 + q2 * 3);
    } else {
      filter4(m, t, w + 1);
    }
  } else if (len == 8) {
    int m = mask_n(t, w, 8);
    int32_t p[4] = {w[3], w[2], w[1], w[0]}, q[4] = {w[4], w[5], w[6], w[7]};
    int fl = flat_n(1, t->shift, p, q, 4);
    if (fl && m) {
      int32_t p3 = w[0], p2 = w[1], p1 = w[2], p0 = w[3], q0 = w[4], q1 = w[5], q2 = w[6], q3 = w[7];
      w[1] = R3(p3 + p3 + p3 + 2 * p2 + p1 + p0 + q0);
      w[2] = R3(p3 + p3 + p2 + 2 * p1 + p0 + q0 + q1);
      w[3] = R3(p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2);
      w[4] = R3(p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3);
      w[5] = R3(p1 + p0 + q0 + 2 * q1 + q2 + q3 + q3);
      w[6] = R3(p0 + q0 + q1 + 2 * q2 + q3 + q3 + q3);
    } else {
      filter4(m, t, w + 2);
    }
  } else {
    int m = mask_n(t, w + 3, 8);
    int32_t p[4] = {w[6], w[5], w[4], w[3]}, q[4] = {w[7], w[8], w[9], w[10]};
    int fl = flat_n(1, t->shift, p, q, 4);
    int32_t p2a[4] = {w[6], w[2], w[1], w[0]}, q2a[4] = {w[7], w[11], w[12], w[13]};
    int fl2 = flat_n(1, t->shift, p2a, q2a, 4);
    if (fl2 && fl && m) {
      int32_t p6 = w[0], p5 = w[1], p4 = w[2], p3 = w[3], p2 = w[4], p1 = w[5], p0 = w[6];
      int32_t q0 = w[7], q1 = w[8], q2 = w[9], q3 = w[10], q4 = w[11], q5 = w[12], q6 = w[13];
      w[1] = R4(p6 * 7 + p5 * 2 + p4 * 2 + p3 + p2 + p1 + p0 + q0);
      w[2] = R4(p6 * 5 + p5 * 2 + p4 * 2 + p3 * 2 + p2 + p1 + p0 + q0 + q1);
      w[3] = R4(p6 * 4 + p5 + p4 * 2 + p3 * 2 + p2 * 2 + p1 + p0 + q0 + q1 + q2);
      w[4] = R4(p6 * 3 + p5 + p4 + p3 * 2 + p2 * 2 + p1 * 2 + p0 + q0 + q1 + q2 + q3);
      w[5] = R4(p6 * 2 + p5 + p4 + p3 + p2 * 2 + p1 * 2 + p0 * 2 + q0 + q1 + q2 + q3 + q4);
      w[6] = R4(p6 + p5 + p4 + p3 + p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + q2 + q3 + q4 + q5);
      w[7] = R4(p5 + p4 + p3 + p2 + p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + q3 + q4 + q5 + q6);
      w[8] = R4(p4 + p3 + p2 + p1 + p0 + q0 * 2 + q1 * 2 + q2 * 2 + q3 + q4 + q5 + q6 * 2);
      w[9] = R4(p3 + p2 + p1 + p0 + q0 + q1 * 2 + q2 * 2 + q3 * 2 + q4 + q5 + q6 * 3);
      w[10] = R4(p2 + p1 + p0 + q0 + q1 + q2 * 2 + q3 * 2 + q4 * 2 + q5 + q6 * 4);
      w[11] = R4(p1 + p0 + q0 + q1 + q2 + q3 * 2 + q4 * 2 + q5 * 2 + q6 * 5);
      w[12] = R4(p0 + q0 + q1 + q2 + q3 + q4 * 2 + q5 * 2 + q6 * 7);
    } else if (fl && m) {
      int32_t p3 = w[3], p2 = w[4], p1 = w[5], p0 = w[6], q0 = w[7], q1 = w[8], q2 = w[9], q3 = w[10];
      w[4] = R3(p3 + p3 + p3 + 2 * p2 + p1 + p0 + q0);
      w[5] = R3(p3 + p3 + p2 + 2 * p1 + p0 + q0 + q1);
      w[6] = R3(p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2);
      w[7] = R3(p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3);
      w[8] = R3(p1 + p0 + q0 + 2 * q1 + q2 + q3 + q3);
      w[9] = R3(p0 + q0 + q1 + 2 * q2 + q3 + q3 + q3);
    } else {
      filter4(m, t, w + 5);
    }
  }
}

static int32_t ld(const void *p, int hbd, long i) { return hbd ? ((const uint16_t *)p)[i] : ((const uint8_t *)p)[i]; }
static void st(void *p, int hbd, long i, int32_t v) {
  if (hbd) ((uint16_t *)p)[i] = (uint16_t)v;
  else ((uint8_t *)p)[i] = (uint8_t)v;
}

/* lpf_{horizontal,vertical}_{4,6,8,14}: 4 lines at sample offset `off`. */
void or_lpf(void *buf, int hbd, int bd, long off, int pitch, int vertical, int len, int level, int sharpness) {
  thr_t t;
  or_lf_thresholds(level, sharpness, &t.mblim, &t.lim, &t.hev);
  t.shift = bd - 8;
  t.bd = bd;
  int32_t w[14];
  for (int i = 0; i < 4; i++) {
    long base = vertical ? off + (long)i * pitch : off + i;
    long step = vertical ? 1 : pitch;
    long start = base - (len / 2) * step;
    for (int k = 0; k < len; k++) w[k] = ld(buf, hbd, start + k * step);
    lpf_line(len, &t, w);
    for (int k = 0; k < len; k++) st(buf, hbd, start + k * step, w[k]);
  }
}

/* enc/deblock.rs:698 lpf_params: (advance in mi, filter length or 0). */
static int lpf_params(const or_deblock_geom *g, int vert, int x, int y, int plane, int ss, int level, int *len) {
  *len = 0;
  if (x >= (g->true_w >> ss) || y >= (g->true_h >> ss)) return 1;
  int mi_row = ss | ((y << ss) >> 2), mi_col = ss | ((x << ss) >> 2);
  long idx = (long)mi_row * g->mi_cols + mi_col;
#define TXW(i) (plane == 0 ? g->tw[i] : ((g->bw[i] >> 1) > 4 ? (g->bw[i] >> 1) : 4))
#define TXH(i) (plane == 0 ? g->th[i] : ((g->bh[i] >> 1) > 4 ? (g->bh[i] >> 1) : 4))
  int ts = vert ? TXW(idx) : TXH(idx);
  int adv = ts >> 2;
  int coord = vert ? x : y;
  if (coord & (ts - 1)) return adv;
  if (coord == 0) return adv;
  long prev = vert ? idx - (1 << ss) : idx - ((long)g->mi_cols << ss);
  int pv = vert ? TXW(prev) : TXH(prev);
  int cs = g->skip ? g->skip[idx] : 0, ps = g->skip ? g->skip[prev] : 0;
  int pu_edge = g->block_id[prev] != g->block_id[idx];
  if (level != 0 && (!ps || !cs || pu_edge)) {
    int md = ts < pv ? ts : pv;
    *len = md == 4 ? 4 : (plane != 0 ? 6 : (md == 8 ? 8 : 14));
  }
#undef TXW
#undef TXH
  return adv;
}

/* enc/deblock.rs:811 filter_plane / :875 filter_plane_hbd. */
void or_deblock_plane(void *buf, int stride, int plane_w, int plane_h, int bd, int plane, int ss, int level_v,
                      int level_h, int sharpness, const or_deblock_geom *g) {
  int hbd = bd > 8, mc = plane_w / 4, mr = plane_h / 4, len;
  for (int my = 0; my < mr; my++)
    for (int mx = 0; mx < mc;) {
      int adv = lpf_params(g, 1, mx * 4, my * 4, plane, ss, level_v, &len);
      if (len) or_lpf(buf, hbd, bd, (long)my * 4 * stride + mx * 4, stride, 1, len, level_v, sharpness);
      mx += adv;
    }
  for (int mx = 0; mx < mc; mx++)
    for (int my = 0; my < mr;) {
      int adv = lpf_params(g, 0, mx * 4, my * 4, plane, ss, level_h, &len);
      if (len) or_lpf(buf, hbd, bd, (long)my * 4 * stride + mx * 4, stride, 0, len, level_h, sharpness);
      my += adv;
    }
}

/* ------------------------------------------------------------------ */
/* CDEF                                                                */
/* ------------------------------------------------------------------ */
#define BSTRIDE 144
#define VB 3
#define HB 8
#define VERY_LARGE 0x7f7f

static int msb(uint32_t n) { return 31 - __builtin_clz(n); }

int or_cdef_constrain(int diff, int threshold, int damping) {
  if (threshold == 0) return 0;
  int shift = damping - msb((uint32_t)threshold);
  if (shift < 0) shift = 0;
  int ad = iabs(diff), m = threshold - (ad >> shift);
  if (m < 0) m = 0;
  int v = ad < m ? ad : m;
  return diff < 0 ? -v : v;
}

int or_cdef_adjust_strength(int strength, int var) {
  int i = (var >> 6) ? msb((uint32_t)(var >> 6)) : 0;
  if (i > 12) i = 12;
  return var ? (strength * (4 + i) + 8) >> 4 : 0;
}

/* dsp/cdef.rs:143 */
int or_cdef_find_dir(const uint16_t *img, int stride, int coeff_shift, int *var_out) {
  static const int DIV[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};
  int32_t cost[8] = {0}, partial[8][15];
  memset(partial, 0, sizeof(partial));
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) {
      int x = (img[i * stride + j] >> coeff_shift) - 128;
      partial[0][i + j] += x;
      partial[1][i + j / 2] += x;
      partial[2][i] += x;
      partial[3][3 + i - j / 2] += x;
      partial[4][7 + i - j] += x;
      partial[5][3 - i / 2 + j] += x;
      partial[6][j] += x;
      partial[7][i / 2 + j] += x;
    }
  for (int i = 0; i < 8; i++) {
    cost[2] += partial[2][i] * partial[2][i];
    cost[6] += partial[6][i] * partial[6][i];
  }
  cost[2] *= DIV[8];
  cost[6] *= DIV[8];
  for (int i = 0; i < 7; i++) {
    cost[0] += (partial[0][i] * partial[0][i] + partial[0][14 - i] * partial[0][14 - i]) * DIV[i + 1];
    cost[4] += (partial[4][i] * partial[4][i] + partial[4][14 - i] * partial[4][14 - i]) * DIV[i + 1];
  }
  cost[0] += partial[0][7] * partial[0][7] * DIV[8];
  cost[4] += partial[4][7] * partial[4][7] * DIV[8];
  for (int i = 1; i < 8; i += 2) {
    for (int j = 0; j < 5; j++) cost[i] += partial[i][3 + j] * partial[i][3 + j];
    cost[i] *= DIV[8];
    for (int j = 0; j < 3; j++)
      cost[i] += (partial[i][j] * partial[i][j] + partial[i][10 - j] * partial[i][10 - j]) * DIV[2 * j + 2];
  }
  int32_t best = 0;
  int bd = 0;
  for (int i = 0; i < 8; i++)
    if (cost[i] > best) { best = cost[i]; bd = i; }
  *var_out = (best - cost[(bd + 4) & 7]) >> 10;
  return bd;
}

/* dsp/cdef.rs:73 direction offsets in the padded buffer */
static int cdef_dir_off(int dir, int k) {
  static const int D[12][2] = {
      {BSTRIDE, 2 * BSTRIDE},      {BSTRIDE, 2 * BSTRIDE - 1}, {-BSTRIDE + 1, -2 * BSTRIDE + 2},
      {1, -BSTRIDE + 2},           {1, 2},                     {1, BSTRIDE + 2},
      {BSTRIDE + 1, 2 * BSTRIDE + 2}, {BSTRIDE, 2 * BSTRIDE + 1}, {BSTRIDE, 2 * BSTRIDE},
      {BSTRIDE, 2 * BSTRIDE - 1},  {-BSTRIDE + 1, -2 * BSTRIDE + 2}, {1, -BSTRIDE + 2}};
  return D[dir + 2][k];
}

/* dsp/cdef.rs:931 / hbd.rs:1765: one 8x8 (or 4x4) block, rows stepped by
 * `sub`, output into dst (dstride). */
void or_cdef_filter_block(int32_t *dst, int dstride, const uint16_t *in, int pri, int sec, int dir, int pdamp,
                          int sdamp, int bw, int bh, int coeff_shift, int sub) {
  static const int PT[2][2] = {{4, 2}, {3, 3}};
  const int *pt = PT[(pri >> coeff_shift) & 1];
  const int stt[2] = {2, 1};
  for (int i = 0; i < bh; i += sub)
    for (int j = 0; j < bw; j++) {
      const uint16_t *c = in + i * BSTRIDE + j;
      int32_t x = (int16_t)c[0];
      int32_t sum = 0, mx = x, mn = x;
      for (int k = 0; k < 2; k++) {
        int dk = cdef_dir_off(dir, k);
        int32_t p0 = (int16_t)c[dk], p1 = (int16_t)c[-dk];
        sum += pt[k] * or_cdef_constrain(p0 - x, pri, pdamp);
        sum += pt[k] * or_cdef_constrain(p1 - x, pri, pdamp);
        if ((uint16_t)p0 != VERY_LARGE && p0 > mx) mx = p0;
        if ((uint16_t)p1 != VERY_LARGE && p1 > mx) mx = p1;
        if (p0 < mn) mn = p0;
        if (p1 < mn) mn = p1;
        int da = cdef_dir_off(dir + 2, k), db = cdef_dir_off(dir - 2, k);
        int32_t s[4] = {(int16_t)c[da], (int16_t)c[-da], (int16_t)c[db], (int16_t)c[-db]};
        for (int m = 0; m < 4; m++) {
          if ((uint16_t)s[m] != VERY_LARGE && s[m] > mx) mx = s[m];
          if (s[m] < mn) mn = s[m];
          sum += stt[k] * or_cdef_constrain(s[m] - x, sec, sdamp);
        }
      }
      int32_t s16 = (int16_t)sum; /* the reference accumulates in wrapping i16 */
      dst[i * dstride + j] = clampi(x + ((8 + s16 - (s16 < 0)) >> 4), mn, mx);
    }
}

/* dsp/cdef.rs:1160 compute_cdef_dist_16bit / :1192 _8bit: SSE of the
 * filtered blocks packed back to back against the source plane, over the
 * dlist (by, bx) blocks of bsize (0 4x4, 1 4x8, 2 8x4, 3 8x8), rows stepped by
 * sub, >> 2 * coeff_shift. */
uint64_t or_cdef_dist(const void *plane, int hbd, long plane_off, int pstride, const void *packed,
                      const uint8_t *dlist, int n, int bsize, int coeff_shift, int sub) {
  const int bw = (bsize == 3 || bsize == 2) ? 8 : 4, bh = (bsize == 3 || bsize == 1) ? 8 : 4;
  uint64_t sum = 0;
  for (int b = 0; b < n; b++) {
    const long p0 = plane_off + (long)dlist[2 * b] * bh * pstride + (long)dlist[2 * b + 1] * bw;
    const long k0 = (long)b * bw * bh;
    for (int i = 0; i < bh; i += sub)
      for (int j = 0; j < bw; j++) {
        const int32_t e = ld(plane, hbd, p0 + (long)i * pstride + j) - ld(packed, hbd, k0 + (long)i * bw + j);
        sum += (uint64_t)(e * e);
      }
  }
  return sum >> (2 * coeff_shift);
}

/* enc/cdef.rs:517 build_src: padded fb source, VERY_LARGE outside the plane. */
static void build_src(uint16_t *src, const void *pre, int hbd, int pw, int ph, int y0, int x0, int vsize, int hsize) {
  for (int r = 0; r < vsize + 2 * VB; r++) {
    int gy = y0 + r - VB;
    for (int c = 0; c < hsize + 2 * HB; c++) {
      int gx = x0 + c - HB;
      src[r * BSTRIDE + c] =
          (gy < 0 || gy >= ph || gx < 0 || gx >= pw) ? VERY_LARGE : (uint16_t)ld(pre, hbd, (long)gy * pw + gx);
    }
  }
}

/* enc/deblock.rs:652 is_8x8_all_skip: all (up to) four mi of the 8x8 whose
 * top-left mi is (mr, mc) carry the signaled skip flag. */
static int all_skip_8x8(const uint8_t *skip, int mi_cols, int mi_rows, int mr, int mc) {
  for (int r = mr; r < mr + 2 && r < mi_rows; r++)
    for (int c = mc; c < mc + 2 && c < mi_cols; c++)
      if (!skip[r * mi_cols + c]) return 0;
  return 1;
}

/* enc/cdef.rs:1166 cdef_search_still / :1374 _hbd for one 64x64 filter
 * block (fbr, fbc), every candidate in fs[0..n_fs).  The 8x8s searched are
 * the fb's dlist (:1202-1210): all of them when skip is NULL, else those
 * that are not all-skip per the per-mi signaled skip map (mi_cols x
 * mi_rows).  Writes mse_y[n_fs] and mse_uv[n_fs] (U + V accumulated; zero
 * rows for an empty dlist); planes are tightly packed (stride = plane
 * width).  chroma may be NULL. */
void or_cdef_search_fb(const void *rec_y, const void *rec_u, const void *rec_v, const void *src_y,
                       const void *src_u, const void *src_v, int width, int height, int bd, int qindex,
                       int subsampling, int fbr, int fbc, const int *fs, int n_fs, uint64_t *mse_y,
                       uint64_t *mse_uv, const uint8_t *skip, int mi_cols, int mi_rows) {
  const int hbd = bd > 8, cs = bd - 8;
  const int damping = 3 + (qindex >> 6);
  const int sub_y = subsampling < 4 ? subsampling : 4, sub_uv = subsampling < 1 ? subsampling : 1;
  int vsize = height - fbr * 64 < 64 ? height - fbr * 64 : 64;
  int hsize = width - fbc * 64 < 64 ? width - fbc * 64 : 64;
  uint16_t *pad = (uint16_t *)malloc(sizeof(uint16_t) * BSTRIDE * (128 + 2 * VB));
  int32_t blk[64];
  int dir[8][8], var[8][8], act[8][8];
  const int base = VB * BSTRIDE + HB;
  for (int by = 0; by < 8; by++)
    for (int bx = 0; bx < 8; bx++)
      act[by][bx] = !skip || !all_skip_8x8(skip, mi_cols, mi_rows, fbr * 16 + by * 2, fbc * 16 + bx * 2);
  build_src(pad, rec_y, hbd, width, height, fbr * 64, fbc * 64, vsize, hsize);
  for (int by = 0; by < vsize / 8; by++)
    for (int bx = 0; bx < hsize / 8; bx++)
      if (act[by][bx]) dir[by][bx] = or_cdef_find_dir(pad + base + by * 8 * BSTRIDE + bx * 8, BSTRIDE, cs, &var[by][bx]);
  for (int g = 0; g < n_fs; g++) {
    int pri = (fs[g] / 4) << cs, s = fs[g] % 4, sec = (s + (s == 3)) << cs;
    int dmp = damping + cs;
    uint64_t d = 0;
    for (int by = 0; by < vsize / 8; by++)
      for (int bx = 0; bx < hsize / 8; bx++) {
        if (!act[by][bx]) continue;
        const uint16_t *in = pad + base + by * 8 * BSTRIDE + bx * 8;
        if (fs[g] == 0) {
          for (int i = 0; i < 8; i += sub_y)
            for (int j = 0; j < 8; j++) blk[i * 8 + j] = in[i * BSTRIDE + j];
        } else {
          int t = or_cdef_adjust_strength(pri, var[by][bx]);
          or_cdef_filter_block(blk, 8, in, t, sec, pri ? dir[by][bx] : 0, dmp, dmp, 8, 8, cs, sub_y);
        }
        for (int i = 0; i < 8; i += sub_y)
          for (int j = 0; j < 8; j++) {
            int32_t e = ld(src_y, hbd, (long)(fbr * 64 + by * 8 + i) * width + fbc * 64 + bx * 8 + j) - blk[i * 8 + j];
            d += (uint64_t)(e * e);
          }
      }
    mse_y[g] = (d >> (2 * cs)) * (uint64_t)sub_y;
  }
  if (rec_u && mse_uv) {
    int cw = width / 2, ch = height / 2;
    for (int g = 0; g < n_fs; g++) mse_uv[g] = 0;
    const void *recs[2] = {rec_u, rec_v}, *srcs[2] = {src_u, src_v};
    for (int pl = 0; pl < 2; pl++) {
      build_src(pad, recs[pl], hbd, cw, ch, fbr * 32, fbc * 32, vsize / 2, hsize / 2);
      for (int g = 0; g < n_fs; g++) {
        int pri = (fs[g] / 4) << cs, s = fs[g] % 4, sec = (s + (s == 3)) << cs;
        int dmp = damping + cs - 1;
        uint64_t d = 0;
        for (int by = 0; by < vsize / 8; by++)
          for (int bx = 0; bx < hsize / 8; bx++) {
            if (!act[by][bx]) continue;
            const uint16_t *in = pad + base + by * 4 * BSTRIDE + bx * 4;
            if (fs[g] == 0) {
              for (int i = 0; i < 4; i += sub_uv)
                for (int j = 0; j < 4; j++) blk[i * 4 + j] = in[i * BSTRIDE + j];
            } else {
              or_cdef_filter_block(blk, 4, in, pri, sec, pri ? dir[by][bx] : 0, dmp, dmp, 4, 4, cs, sub_uv);
            }
            for (int i = 0; i < 4; i += sub_uv)
              for (int j = 0; j < 4; j++) {
                int32_t e = ld(srcs[pl], hbd, (long)(fbr * 32 + by * 4 + i) * cw + fbc * 32 + bx * 4 + j) -
                            blk[i * 4 + j];
                d += (uint64_t)(e * e);
              }
          }
        mse_uv[g] += (d >> (2 * cs)) * (uint64_t)sub_uv;
      }
    }
  }
  free(pad);
}

/* enc/cdef.rs:343 apply_cdef_frame / :556 apply_cdef_frame_hbd: the
 * decoder-exact CDEF frame pass, out of place (every read is of the
 * post-deblock, pre-CDEF planes in_*; every pixel of out_* is written: the
 * filtered value, or the input where nothing filters).  Per 64x64 fb
 * (raster): the dlist of 8x8s not all-skip (skip NULL: every 8x8); none ->
 * the fb is copied.  Strength index per fb fb_str[2 fb] (luma),
 * [2 fb + 1] (chroma): primary idx / 4, secondary idx % 4 with 3 -> 4, both
 * << coeff_shift; damping + coeff_shift (chroma one less).  Luma: the
 * direction search always runs (chroma reuses it); a block filters when its
 * variance-adjusted primary or the secondary is nonzero, along its
 * direction when the fb's primary is nonzero, else direction 0.  Chroma
 * (in_u non-NULL, 4:2:0): 4x4 blocks of the same dlist, unadjusted primary.
 * stats (may be NULL): [0] pixels of every filter invocation, [1] of those
 * the pixels that changed. */
void or_cdef_apply_frame(const void *in_y, const void *in_u, const void *in_v, void *out_y, void *out_u,
                         void *out_v, int width, int height, int bd, int damping, const uint8_t *fb_str,
                         const uint8_t *skip, int mi_cols, int mi_rows, uint64_t *stats) {
  const int hbd = bd > 8, cs = bd - 8;
  const int nhfb = (width + 63) / 64, nvfb = (height + 63) / 64;
  const int cw = width / 2, ch = height / 2;
  uint16_t *pad = (uint16_t *)malloc(sizeof(uint16_t) * BSTRIDE * (64 + 2 * VB));
  const int base = VB * BSTRIDE + HB;
  int32_t blk[64];
  uint64_t filtered = 0, changed = 0;
  /* start from the input: positions nothing filters keep it */
  for (long i = 0; i < (long)width * height; i++) st(out_y, hbd, i, ld(in_y, hbd, i));
  if (in_u)
    for (long i = 0; i < (long)cw * ch; i++) st(out_u, hbd, i, ld(in_u, hbd, i)), st(out_v, hbd, i, ld(in_v, hbd, i));
  for (int fbr = 0; fbr < nvfb; fbr++)
    for (int fbc = 0; fbc < nhfb; fbc++) {
      const int vsize = height - fbr * 64 < 64 ? height - fbr * 64 : 64;
      const int hsize = width - fbc * 64 < 64 ? width - fbc * 64 : 64;
      int act[8][8], dir[8][8], var[8][8], n_act = 0;
      for (int by = 0; by < 8; by++)
        for (int bx = 0; bx < 8; bx++) {
          act[by][bx] = by < vsize / 8 && bx < hsize / 8 &&
                        (!skip || !all_skip_8x8(skip, mi_cols, mi_rows, fbr * 16 + by * 2, fbc * 16 + bx * 2));
          n_act += act[by][bx];
        }
      if (!n_act) continue;
      const int fb = fbr * nhfb + fbc, sy = fb_str[2 * fb], suv = fb_str[2 * fb + 1];
      const int lvl_y = (sy / 4) << cs, sec_y = ((sy % 4) + (sy % 4 == 3)) << cs;
      const int lvl_uv = in_u ? (suv / 4) << cs : 0, sec_uv = in_u ? ((suv % 4) + (suv % 4 == 3)) << cs : 0;
      const int dmp = damping + cs;
      build_src(pad, in_y, hbd, width, height, fbr * 64, fbc * 64, vsize, hsize);
      for (int by = 0; by < 8; by++)
        for (int bx = 0; bx < 8; bx++)
          if (act[by][bx]) dir[by][bx] = or_cdef_find_dir(pad + base + by * 8 * BSTRIDE + bx * 8, BSTRIDE, cs, &var[by][bx]);
      if (lvl_y || sec_y)
        for (int by = 0; by < 8; by++)
          for (int bx = 0; bx < 8; bx++) {
            if (!act[by][bx]) continue;
            const int t = or_cdef_adjust_strength(lvl_y, var[by][bx]);
            if (t == 0 && sec_y == 0) continue;
            const uint16_t *in = pad + base + by * 8 * BSTRIDE + bx * 8;
            or_cdef_filter_block(blk, 8, in, t, sec_y, lvl_y ? dir[by][bx] : 0, dmp, dmp, 8, 8, cs, 1);
            for (int i = 0; i < 8; i++)
              for (int j = 0; j < 8; j++) {
                const long o = (long)(fbr * 64 + by * 8 + i) * width + fbc * 64 + bx * 8 + j;
                changed += blk[i * 8 + j] != (int32_t)in[i * BSTRIDE + j];
                st(out_y, hbd, o, blk[i * 8 + j]);
              }
            filtered += 64;
          }
      if (in_u && (lvl_uv || sec_uv)) {
        const void *ins[2] = {in_u, in_v};
        void *outs[2] = {out_u, out_v};
        for (int pl = 0; pl < 2; pl++) {
          build_src(pad, ins[pl], hbd, cw, ch, fbr * 32, fbc * 32, vsize / 2, hsize / 2);
          for (int by = 0; by < 8; by++)
            for (int bx = 0; bx < 8; bx++) {
              if (!act[by][bx]) continue;
              const uint16_t *in = pad + base + by * 4 * BSTRIDE + bx * 4;
              or_cdef_filter_block(blk, 4, in, lvl_uv, sec_uv, lvl_uv ? dir[by][bx] : 0, dmp - 1, dmp - 1, 4, 4, cs,
                                   1);
              for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                  const long o = (long)(fbr * 32 + by * 4 + i) * cw + fbc * 32 + bx * 4 + j;
                  changed += blk[i * 4 + j] != (int32_t)in[i * BSTRIDE + j];
                  st(outs[pl], hbd, o, blk[i * 4 + j]);
                }
              filtered += 16;
            }
        }
      }
    }
  free(pad);
  if (stats) stats[0] = filtered, stats[1] = changed;
}

/* enc/deblock.rs:208 plane_sse + :228 try_filter_plane: SSE of the source
 * against a scratch copy of the recon filtered at (level, level). */
static int64_t try_level(const void *recon, const void *src, int pw, int ph, int bd, int plane, int ss, int sharpness,
                         const or_deblock_geom *g, int level) {
  const int hbd = bd > 8;
  const size_t bytes = (size_t)pw * ph * (hbd ? 2 : 1);
  void *scratch = malloc(bytes);
  memcpy(scratch, recon, bytes);
  if (level) or_deblock_plane(scratch, pw, pw, ph, bd, plane, ss, level, level, sharpness, g);
  uint64_t s = 0;
  for (long i = 0; i < (long)pw * ph; i++) {
    const int64_t d = (int64_t)ld(src, hbd, i) - ld(scratch, hbd, i);
    s += (uint64_t)(d * d);
  }
  free(scratch);
  return (int64_t)s;
}

/* enc/deblock.rs:337 search_filter_level on the still path (last levels 0,
 * luma dir 2 / chroma dir 0: one level for both edge directions), with a
 * real trial filter per evaluated level.  Planes tight (stride = width). */
void or_deblock_search_level(const void *recon, const void *src, int pw, int ph, int bd, int plane, int ss,
                             int sharpness, const or_deblock_geom *g, int conv_th, int *best, int64_t *err0,
                             int64_t *best_err) {
  const int maxl = 63;
  int64_t ss_err[64];
  for (int i = 0; i < 64; i++) ss_err[i] = -1;
  int filt_mid = 0, filt_dir = 0;
  int step = filt_mid < 16 ? 4 : filt_mid / 4;
  int64_t be = try_level(recon, src, pw, ph, bd, plane, ss, sharpness, g, filt_mid);
  int fb = filt_mid;
  ss_err[filt_mid] = be;
  int conv = 0;
  while (step > 0) {
    const int hi = filt_mid + step > maxl ? maxl : filt_mid + step;
    const int lo = filt_mid - step < 0 ? 0 : filt_mid - step;
    int64_t bias = (be >> (15 - (filt_mid / 8))) * (int64_t)step;
    bias >>= 1;
    if (filt_dir <= 0 && lo != filt_mid) {
      if (ss_err[lo] < 0) ss_err[lo] = try_level(recon, src, pw, ph, bd, plane, ss, sharpness, g, lo);
      if (ss_err[lo] < be + bias) {
        if (ss_err[lo] < be) be = ss_err[lo];
        fb = lo;
      }
    }
    if (filt_dir >= 0 && hi != filt_mid) {
      if (ss_err[hi] < 0) ss_err[hi] = try_level(recon, src, pw, ph, bd, plane, ss, sharpness, g, hi);
      if (ss_err[hi] < be - bias) {
        be = ss_err[hi];
        fb = hi;
      }
    }
    if (fb == filt_mid) {
      conv++;
      if (conv == conv_th) step = 0;
      else step /= 2;
      filt_dir = 0;
    } else {
      filt_dir = fb < filt_mid ? -1 : 1;
      filt_mid = fb;
    }
  }
  *best = fb;
  *err0 = ss_err[0];
  *best_err = ss_err[fb];
}
