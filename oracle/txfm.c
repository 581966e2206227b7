/*
 * Oracle: forward and inverse AV1 transforms exactly as zenav1-svt computes
 * them (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h).
 *
 * The DCT butterfly networks are written as the recursion they are: an N-point
 * DCT is a mirror add/sub stage, an N/2-point DCT on the sums (even outputs)
 * and an "odd part" network on the differences (odd outputs).  The recursion
 * reproduces, operation for operation and rounding point for rounding point,
 * the explicit stage lists of
 *   fdct4/8/16/32/64   dsp/fwd_txfm.rs:184,208,274,405,729
 *   idct4/8/16/32/64   dsp/inv_txfm.rs:79,106,259,402,739
 * (a one-off symbolic expansion of both forms gave identical expression trees
 * for every output; the golden vectors in tests/ pin the numbers).
 * ADST networks follow fadst4/8/16 (dsp/fwd_txfm.rs:1497,1565,1651) and
 * iadst4/8/16 (dsp/inv_txfm.rs:177,1497,1581) stage by stage.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "zav1_oracle.h"

/* dsp/fwd_txfm.rs:22 COSPI_ARR_DATA = round(cos(j*pi/128) * 2^(10+i)). */
static int32_t g_cospi[7][64];
static int g_cospi_ready;
/* dsp/fwd_txfm.rs:56 SINPI_ARR_DATA (libaom sinpi rows; data, not a closed form). */
static const int32_t g_sinpi[7][5] = {
    {0, 330, 621, 836, 951},         {0, 660, 1241, 1672, 1901},
    {0, 1321, 2482, 3344, 3803},     {0, 2642, 4964, 6689, 7606},
    {0, 5283, 9929, 13377, 15212},   {0, 10566, 19858, 26755, 30424},
    {0, 21133, 39716, 53510, 60849},
};

static void init_cospi(void) {
  if (g_cospi_ready) return;
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < 64; j++)
      g_cospi[i][j] = (int32_t)floor(cos(j * M_PI / 128.0) * (double)(1 << (10 + i)) + 0.5);
  g_cospi_ready = 1;
}

const int32_t *or_cospi_row(int cos_bit) { init_cospi(); return g_cospi[cos_bit - 10]; }
const int32_t *or_sinpi_row(int cos_bit) { return g_sinpi[cos_bit - 10]; }

/* dsp/fwd_txfm.rs:156 half_btf: i64 products, round-shift by cos_bit. */
static int32_t btf(int32_t w0, int32_t a, int32_t w1, int32_t b, int cb) {
  int64_t r = (int64_t)w0 * a + (int64_t)w1 * b;
  return (int32_t)((r + ((int64_t)1 << (cb - 1))) >> cb);
}

/* dsp/fwd_txfm.rs:146 round_shift_i64 */
static int32_t rshift64(int64_t v, int bit) {
  if (bit == 0) return (int32_t)v;
  return (int32_t)((v + ((int64_t)1 << (bit - 1))) >> bit);
}

/* dsp/inv_txfm.rs:27 clamp_value (bit <= 0 is a no-op). */
static int32_t clampv(int32_t v, int bit) {
  if (bit <= 0) return v;
  int64_t mx = ((int64_t)1 << (bit - 1)) - 1, mn = -((int64_t)1 << (bit - 1));
  return (int32_t)(v < mn ? mn : (v > mx ? mx : v));
}

static int bitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
  return r;
}
static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

/* ------------------------------------------------------------------ */
/* forward DCT                                                         */
/* ------------------------------------------------------------------ */
typedef struct { const int32_t *c; int cb; } ctx_t;
#define C(k) (cx->c[(k) < 0 ? -(k) : (k)] * ((k) < 0 ? -1 : 1))

/* Odd part of the forward DCT on M differences u[] (u[j] = x[N/2-1-j] -
 * x[N/2+j]); writes the odd DCT outputs to out[oidx[k]] (k = natural
 * odd-coefficient index). */
static void fodd(const ctx_t *cx, int32_t *u, int M, int32_t *out, const int *oidx) {
  int lg = ilog2(M);
  if (M >= 4) {
    for (int j = M / 4; j < M / 2; j++) {
      int m = M - 1 - j;
      int32_t nj = btf(-cx->c[32], u[j], cx->c[32], u[m], cx->cb);
      int32_t nm = btf(cx->c[32], u[m], cx->c[32], u[j], cx->cb);
      u[j] = nj; u[m] = nm;
    }
  }
  for (int S = M / 2; S >= 2; S /= 2) {
    for (int k = 0; k < M / S; k++) {
      int p = k * S;
      for (int i = 0; i < S / 2; i++) {
        int lo = p + i, hi = p + S - 1 - i;
        int32_t a = u[lo], b = u[hi];
        if ((k & 1) == 0) { u[lo] = a + b; u[hi] = a - b; }
        else              { u[lo] = b - a; u[hi] = b + a; }
      }
    }
    if (S >= 4) {
      int nbh = (M / 2) / S, bits = ilog2(nbh);
      for (int k = 0; k < nbh; k++) {
        int p = k * S;
        int A = (32 * S / M) * (1 + 4 * bitrev(k, bits)), B = 64 - A;
        for (int j = p + S / 4; j < p + S / 2; j++) {
          int m = M - 1 - j;
          int32_t nj = btf(-cx->c[A], u[j], cx->c[B], u[m], cx->cb);
          int32_t nm = btf(cx->c[A], u[m], cx->c[B], u[j], cx->cb);
          u[j] = nj; u[m] = nm;
        }
        for (int j = p + S / 2; j < p + 3 * S / 4; j++) {
          int m = M - 1 - j;
          int32_t nj = btf(-cx->c[B], u[j], -cx->c[A], u[m], cx->cb);
          int32_t nm = btf(cx->c[B], u[m], -cx->c[A], u[j], cx->cb);
          u[j] = nj; u[m] = nm;
        }
      }
    }
  }
  for (int j = 0; j < M / 2; j++) {
    int m = M - 1 - j;
    int Q = (32 / M) * (1 + 4 * bitrev(j, lg - 1));
    int32_t nj = btf(cx->c[64 - Q], u[j], cx->c[Q], u[m], cx->cb);
    int32_t nm = btf(cx->c[64 - Q], u[m], -cx->c[Q], u[j], cx->cb);
    u[j] = nj; u[m] = nm;
  }
  for (int j = 0; j < M; j++) out[oidx[bitrev(j, lg)]] = u[j];
}

static void fdct_rec(const ctx_t *cx, const int32_t *v, int N, int32_t *out, const int *oidx) {
  if (N == 2) {
    out[oidx[0]] = btf(cx->c[32], v[0], cx->c[32], v[1], cx->cb);
    out[oidx[1]] = btf(-cx->c[32], v[1], cx->c[32], v[0], cx->cb);
    return;
  }
  int M = N / 2;
  int32_t a[32], b[32];
  int ei[32], oi[32];
  for (int i = 0; i < M; i++) {
    a[i] = v[i] + v[N - 1 - i];
    b[M - 1 - i] = v[i] - v[N - 1 - i];
    ei[i] = oidx[2 * i];
    oi[i] = oidx[2 * i + 1];
  }
  fdct_rec(cx, a, M, out, ei);
  fodd(cx, b, M, out, oi);
}

static void fdct(const int32_t *in, int32_t *out, int n, int cos_bit) {
  ctx_t cx = {or_cospi_row(cos_bit), cos_bit};
  int idx[64];
  for (int i = 0; i < n; i++) idx[i] = i;
  fdct_rec(&cx, in, n, out, idx);
}

/* ------------------------------------------------------------------ */
/* forward ADST / identity                                             */
/* ------------------------------------------------------------------ */
/* dsp/fwd_txfm.rs:1497 fadst4 (i64 intermediates, all-zero early out). */
static void fadst4(const int32_t *in, int32_t *out, int cos_bit) {
  const int32_t *s = g_sinpi[cos_bit - 10];
  int32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
  if ((x0 | x1 | x2 | x3) == 0) { out[0] = out[1] = out[2] = out[3] = 0; return; }
  int64_t s0 = (int64_t)s[1] * x0, s1 = (int64_t)s[4] * x0, s2 = (int64_t)s[2] * x1;
  int64_t s3 = (int64_t)s[1] * x1, s4 = (int64_t)s[3] * x2, s5 = (int64_t)s[4] * x3;
  int64_t s6 = (int64_t)s[2] * x3;
  int64_t s7 = (int64_t)(int32_t)(x0 + x1) - x3;
  int64_t y0 = s0 + s2 + s5;
  int64_t y1 = (int64_t)s[3] * s7;
  int64_t y2 = s1 - s3 + s6;
  int64_t y3 = s4;
  out[0] = rshift64(y0 + y3, cos_bit);
  out[1] = rshift64(y1, cos_bit);
  out[2] = rshift64(y2 - y3, cos_bit);
  out[3] = rshift64(y2 - y0 + y3, cos_bit);
}

/* Shared stage helpers for the 8/16-point ADST networks. */
static void add_sub_dist(int32_t *v, int n, int d, int clamp_bit) {
  /* blocks of 2d: v[b+i] = v[b+i] + v[b+i+d], v[b+i+d] = v[b+i] - v[b+i+d] */
  for (int b = 0; b < n; b += 2 * d)
    for (int i = 0; i < d; i++) {
      int32_t x = v[b + i], y = v[b + i + d];
      v[b + i] = clampv(x + y, clamp_bit);
      v[b + i + d] = clampv(x - y, clamp_bit);
    }
}
/* pair rotation on (p, p+1) with angle X: form 1 and form 2 of the ADST
 * networks (fadst8 stage 4, fadst16 stages 4/6; iadst mirrors). */
static void rot_f1(int32_t *v, int p, int X, const int32_t *c, int cb) {
  int32_t a = v[p], b = v[p + 1];
  v[p] = btf(c[X], a, c[64 - X], b, cb);
  v[p + 1] = btf(c[64 - X], a, -c[X], b, cb);
}
static void rot_f2(int32_t *v, int p, int X, const int32_t *c, int cb) {
  int32_t a = v[p], b = v[p + 1];
  v[p] = btf(-c[64 - X], a, c[X], b, cb);
  v[p + 1] = btf(c[X], a, c[64 - X], b, cb);
}
static void rot32(int32_t *v, int p, const int32_t *c, int cb) {
  int32_t a = v[p], b = v[p + 1];
  v[p] = btf(c[32], a, c[32], b, cb);
  v[p + 1] = btf(c[32], a, -c[32], b, cb);
}

/* dsp/fwd_txfm.rs:1565 fadst8 */
static void fadst8(const int32_t *in, int32_t *out, int cos_bit) {
  const int32_t *c = or_cospi_row(cos_bit);
  int32_t v[8] = {in[0], -in[7], -in[3], in[4], -in[1], in[6], in[2], -in[5]};
  rot32(v, 2, c, cos_bit); rot32(v, 6, c, cos_bit);
  add_sub_dist(v, 8, 2, 0);
  rot_f1(v, 4, 16, c, cos_bit); rot_f2(v, 6, 16, c, cos_bit);
  add_sub_dist(v, 8, 4, 0);
  for (int k = 0; k < 4; k++) rot_f1(v, 2 * k, 4 + 16 * k, c, cos_bit);
  static const int perm[8] = {1, 6, 3, 4, 5, 2, 7, 0};
  for (int i = 0; i < 8; i++) out[i] = v[perm[i]];
}

/* dsp/fwd_txfm.rs:1651 fadst16 */
static void fadst16(const int32_t *in, int32_t *out, int cos_bit) {
  const int32_t *c = or_cospi_row(cos_bit);
  int32_t v[16] = {in[0], -in[15], -in[7], in[8], -in[3], in[12], in[4], -in[11],
                   -in[1], in[14], in[6], -in[9], in[2], -in[13], -in[5], in[10]};
  for (int p = 2; p < 16; p += 4) rot32(v, p, c, cos_bit);
  add_sub_dist(v, 16, 2, 0);
  for (int b = 0; b < 16; b += 8) { rot_f1(v, b + 4, 16, c, cos_bit); rot_f2(v, b + 6, 16, c, cos_bit); }
  add_sub_dist(v, 16, 4, 0);
  rot_f1(v, 8, 8, c, cos_bit); rot_f1(v, 10, 40, c, cos_bit);
  rot_f2(v, 12, 8, c, cos_bit); rot_f2(v, 14, 40, c, cos_bit);
  add_sub_dist(v, 16, 8, 0);
  for (int k = 0; k < 8; k++) rot_f1(v, 2 * k, 2 + 8 * k, c, cos_bit);
  static const int perm[16] = {1, 14, 3, 12, 5, 10, 7, 8, 9, 6, 11, 4, 13, 2, 15, 0};
  for (int i = 0; i < 16; i++) out[i] = v[perm[i]];
}

/* dsp/fwd_txfm.rs:1553,1830,1840,1484,1474 fidentity4..64 */
static void fidentity(const int32_t *in, int32_t *out, int n) {
  for (int i = 0; i < n; i++) {
    switch (n) {
      case 4: out[i] = rshift64((int64_t)in[i] * 5793, 12); break;
      case 8: out[i] = in[i] * 2; break;
      case 16: out[i] = rshift64((int64_t)in[i] * 2 * 5793, 12); break;
      case 32: out[i] = in[i] * 4; break;
      default: out[i] = rshift64((int64_t)in[i] * 4 * 5793, 12); break;
    }
  }
}

/* dsp/fwd_txfm.rs:1855 get_fwd_txfm_func: kind 0 DCT, 1/2 ADST, 3 identity. */
int or_fwd_txfm1d(int kind, int n, const int32_t *in, int32_t *out, int cos_bit) {
  init_cospi();
  if (kind == 0) {
    if (n < 4 || n > 64) return 0;
    fdct(in, out, n, cos_bit);
    return 1;
  }
  if (kind == 1 || kind == 2) {
    if (n == 4) fadst4(in, out, cos_bit);
    else if (n == 8) fadst8(in, out, cos_bit);
    else if (n == 16) fadst16(in, out, cos_bit);
    else return 0;
    return 1;
  }
  if (kind == 3) { fidentity(in, out, n); return 1; }
  return 0;
}

/* ------------------------------------------------------------------ */
/* inverse DCT (INV_COS_BIT 12, clamps on every add/sub)               */
/* ------------------------------------------------------------------ */
static void iodd(const int32_t *c, const int32_t *coef, int M, int32_t *u, int rng) {
  int lg = ilog2(M);
  const int cb = 12;
  for (int j = 0; j < M; j++) u[j] = coef[bitrev(j, lg)];
  for (int j = 0; j < M / 2; j++) {
    int m = M - 1 - j;
    int Q = (32 / M) * (1 + 4 * bitrev(j, lg - 1));
    int32_t nj = btf(c[64 - Q], u[j], -c[Q], u[m], cb);
    int32_t nm = btf(c[Q], u[j], c[64 - Q], u[m], cb);
    u[j] = nj; u[m] = nm;
  }
  for (int S = 2; S <= M / 2; S *= 2) {
    if (S >= 4) {
      int nbh = (M / 2) / S, bits = ilog2(nbh);
      for (int k = 0; k < nbh; k++) {
        int p = k * S;
        int A = (32 * S / M) * (1 + 4 * bitrev(k, bits)), B = 64 - A;
        for (int j = p + S / 4; j < p + S / 2; j++) {
          int m = M - 1 - j;
          int32_t nj = btf(-c[A], u[j], c[B], u[m], cb);
          int32_t nm = btf(c[B], u[j], c[A], u[m], cb);
          u[j] = nj; u[m] = nm;
        }
        for (int j = p + S / 2; j < p + 3 * S / 4; j++) {
          int m = M - 1 - j;
          int32_t nj = btf(-c[B], u[j], -c[A], u[m], cb);
          int32_t nm = btf(-c[A], u[j], c[B], u[m], cb);
          u[j] = nj; u[m] = nm;
        }
      }
    }
    for (int k = 0; k < M / S; k++) {
      int p = k * S;
      for (int i = 0; i < S / 2; i++) {
        int lo = p + i, hi = p + S - 1 - i;
        int32_t a = u[lo], b = u[hi];
        if ((k & 1) == 0) { u[lo] = clampv(a + b, rng); u[hi] = clampv(a - b, rng); }
        else              { u[lo] = clampv(b - a, rng); u[hi] = clampv(b + a, rng); }
      }
    }
  }
  if (M >= 4) {
    for (int j = M / 4; j < M / 2; j++) {
      int m = M - 1 - j;
      int32_t nj = btf(-c[32], u[j], c[32], u[m], cb);
      int32_t nm = btf(c[32], u[j], c[32], u[m], cb);
      u[j] = nj; u[m] = nm;
    }
  }
}

static void idct_rec(const int32_t *c, const int32_t *coef, int N, int32_t *y, int rng) {
  if (N == 2) {
    y[0] = btf(c[32], coef[0], c[32], coef[1], 12);
    y[1] = btf(c[32], coef[0], -c[32], coef[1], 12);
    return;
  }
  int M = N / 2;
  int32_t ev[32], od[32], E[32], u[32];
  for (int k = 0; k < M; k++) { ev[k] = coef[2 * k]; od[k] = coef[2 * k + 1]; }
  idct_rec(c, ev, M, E, rng);
  iodd(c, od, M, u, rng);
  for (int i = 0; i < M; i++) {
    y[i] = clampv(E[i] + u[M - 1 - i], rng);
    y[N - 1 - i] = clampv(E[i] - u[M - 1 - i], rng);
  }
}

/* dsp/inv_txfm.rs:177 iadst4 (i32 products as the reference computes them). */
static void iadst4(const int32_t *in, int32_t *out) {
  const int32_t *s = g_sinpi[2];
  int32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
  if ((x0 | x1 | x2 | x3) == 0) { out[0] = out[1] = out[2] = out[3] = 0; return; }
  int32_t s0 = s[1] * x0, s1 = s[2] * x0, s2 = s[3] * x1, s3 = s[4] * x2;
  int32_t s4 = s[1] * x2, s5 = s[2] * x3, s6 = s[4] * x3;
  int32_t s7 = (x0 - x2) + x3;
  s0 = s0 + s3;
  s1 = s1 - s4;
  s3 = s2;
  s2 = s[3] * s7;
  s0 = s0 + s5;
  s1 = s1 - s6;
  int32_t y0 = s0 + s3, y1 = s1 + s3, y2 = s2, y3 = s0 + s1;
  y3 = y3 - s3;
  out[0] = rshift64(y0, 12); out[1] = rshift64(y1, 12);
  out[2] = rshift64(y2, 12); out[3] = rshift64(y3, 12);
}

/* dsp/inv_txfm.rs:1497 iadst8 */
static void iadst8(const int32_t *in, int32_t *out, int rng) {
  const int32_t *c = or_cospi_row(12);
  int32_t v[8] = {in[7], in[0], in[5], in[2], in[3], in[4], in[1], in[6]};
  for (int k = 0; k < 4; k++) rot_f1(v, 2 * k, 4 + 16 * k, c, 12);
  add_sub_dist(v, 8, 4, rng);
  rot_f1(v, 4, 16, c, 12); rot_f2(v, 6, 16, c, 12);
  add_sub_dist(v, 8, 2, rng);
  rot32(v, 2, c, 12); rot32(v, 6, c, 12);
  out[0] = v[0]; out[1] = -v[4]; out[2] = v[6]; out[3] = -v[2];
  out[4] = v[3]; out[5] = -v[7]; out[6] = v[5]; out[7] = -v[1];
}

/* dsp/inv_txfm.rs:1581 iadst16 */
static void iadst16(const int32_t *in, int32_t *out, int rng) {
  const int32_t *c = or_cospi_row(12);
  int32_t v[16] = {in[15], in[0], in[13], in[2], in[11], in[4], in[9], in[6],
                   in[7], in[8], in[5], in[10], in[3], in[12], in[1], in[14]};
  for (int k = 0; k < 8; k++) rot_f1(v, 2 * k, 2 + 8 * k, c, 12);
  add_sub_dist(v, 16, 8, rng);
  rot_f1(v, 8, 8, c, 12); rot_f1(v, 10, 40, c, 12);
  rot_f2(v, 12, 8, c, 12); rot_f2(v, 14, 40, c, 12);
  add_sub_dist(v, 16, 4, rng);
  for (int b = 0; b < 16; b += 8) { rot_f1(v, b + 4, 16, c, 12); rot_f2(v, b + 6, 16, c, 12); }
  add_sub_dist(v, 16, 2, rng);
  for (int p = 2; p < 16; p += 4) rot32(v, p, c, 12);
  static const int src[16] = {0, 8, 12, 4, 6, 14, 10, 2, 3, 11, 15, 7, 5, 13, 9, 1};
  for (int i = 0; i < 16; i++) out[i] = (i & 1) ? -v[src[i]] : v[src[i]];
}

/* dsp/inv_txfm.rs:236,247,1760,727,1486 iidentity4..64 */
static void iidentity(const int32_t *in, int32_t *out, int n) { fidentity(in, out, n); }

int or_inv_txfm1d(int kind, int n, const int32_t *in, int32_t *out, int range) {
  init_cospi();
  if (kind == 0) {
    if (n < 4 || n > 64) return 0;
    idct_rec(or_cospi_row(12), in, n, out, range);
    return 1;
  }
  if (kind == 1 || kind == 2) {
    if (n == 4) iadst4(in, out);
    else if (n == 8) iadst8(in, out, range);
    else if (n == 16) iadst16(in, out, range);
    else return 0;
    return 1;
  }
  if (kind == 3) { iidentity(in, out, n); return 1; }
  return 0;
}

/* ------------------------------------------------------------------ */
/* 2D                                                                  */
/* ------------------------------------------------------------------ */
void or_tx_size_dims(int tx_size, int *w, int *h) {
  static const int W[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
  static const int H[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};
  *w = W[tx_size];
  *h = H[tx_size];
}

/* dsp/txfm_dispatch.rs:89 flip_cfg / :99 tx_type_to_1d, TxType C order. */
static void tx_type_cfg(int tx_type, int *col, int *row, int *ud, int *lr) {
  static const int COL[16] = {0, 1, 0, 1, 2, 0, 2, 1, 2, 3, 0, 3, 1, 3, 2, 3};
  static const int ROW[16] = {0, 0, 1, 1, 0, 2, 2, 2, 1, 3, 3, 0, 3, 1, 3, 2};
  *col = COL[tx_type];
  *row = ROW[tx_type];
  *ud = (tx_type == 4 || tx_type == 8 || tx_type == 14 || tx_type == 6);
  *lr = (tx_type == 5 || tx_type == 7 || tx_type == 15 || tx_type == 6);
}

/* dsp/fwd_txfm.rs:91/:101 FWD_COS_BIT_COL/ROW, :110 fwd_txfm_shift */
static const int8_t FWD_CB_COL[5][5] = {{13, 13, 13, 0, 0}, {13, 13, 13, 12, 0}, {13, 13, 13, 12, 13},
                                        {0, 13, 13, 12, 13}, {0, 0, 13, 12, 13}};
static const int8_t FWD_CB_ROW[5][5] = {{13, 13, 12, 0, 0}, {13, 13, 13, 12, 0}, {13, 13, 12, 13, 12},
                                        {0, 12, 13, 12, 11}, {0, 0, 12, 11, 10}};
static void fwd_shift(int w, int h, int s[3]) {
  int lw = ilog2(w), lh = ilog2(h);
  int key = lw * 8 + lh;
  switch (key) {
    case 2 * 8 + 2: s[0] = 2; s[1] = 0; s[2] = 0; break;
    case 3 * 8 + 3: s[0] = 2; s[1] = -1; s[2] = 0; break;
    case 4 * 8 + 4: s[0] = 2; s[1] = -2; s[2] = 0; break;
    case 5 * 8 + 5: s[0] = 2; s[1] = -4; s[2] = 0; break;
    case 6 * 8 + 6: s[0] = 0; s[1] = -2; s[2] = -2; break;
    case 2 * 8 + 3: case 3 * 8 + 2: s[0] = 2; s[1] = -1; s[2] = 0; break;
    case 3 * 8 + 4: case 4 * 8 + 3: s[0] = 2; s[1] = -2; s[2] = 0; break;
    case 4 * 8 + 5: case 5 * 8 + 4: s[0] = 2; s[1] = -4; s[2] = 0; break;
    case 5 * 8 + 6: s[0] = 0; s[1] = -2; s[2] = -2; break;
    case 6 * 8 + 5: s[0] = 2; s[1] = -4; s[2] = -2; break;
    case 2 * 8 + 4: case 4 * 8 + 2: s[0] = 2; s[1] = -1; s[2] = 0; break;
    case 3 * 8 + 5: case 5 * 8 + 3: s[0] = 2; s[1] = -2; s[2] = 0; break;
    case 4 * 8 + 6: s[0] = 0; s[1] = -2; s[2] = 0; break;
    case 6 * 8 + 4: s[0] = 2; s[1] = -4; s[2] = 0; break;
    default: s[0] = s[1] = s[2] = 0; break;
  }
}

/* dsp/fwd_txfm.rs:126 round_shift_array: bit>0 rounds right, bit<0 shifts left. */
static void rs_array(int32_t *a, int n, int bit) {
  if (bit == 0) return;
  for (int i = 0; i < n; i++) {
    if (bit > 0) a[i] = (a[i] + (1 << (bit - 1))) >> bit;
    else a[i] = (int32_t)((uint32_t)a[i] << (-bit));
  }
}

static int kind_ok(int kind, int n) { return !((kind == 1 || kind == 2) && n > 16); }

/* dsp/fwd_txfm.rs:1891 fwd_txfm2d_core + :1959 fwd_txfm2d_c_exact */
int or_fwd_txfm2d(const int32_t *in, int32_t *out, int stride, int tx_size, int tx_type) {
  int w, h, col, row, ud, lr, sh[3];
  if (tx_size < 0 || tx_size >= OR_TX_SIZES_ALL || tx_type < 0 || tx_type > 15) return 0;
  or_tx_size_dims(tx_size, &w, &h);
  tx_type_cfg(tx_type, &col, &row, &ud, &lr);
  if (!kind_ok(col, h) || !kind_ok(row, w)) return 0;
  init_cospi();
  fwd_shift(w, h, sh);
  int cbc = FWD_CB_COL[ilog2(w) - 2][ilog2(h) - 2], cbr = FWD_CB_ROW[ilog2(w) - 2][ilog2(h) - 2];
  int rect = abs(ilog2(w) - ilog2(h)) == 1;
  int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * w * h);
  int32_t tin[64], tout[64];
  for (int c = 0; c < w; c++) {
    for (int r = 0; r < h; r++) tin[r] = in[(ud ? (h - 1 - r) : r) * stride + c];
    rs_array(tin, h, -sh[0]);
    or_fwd_txfm1d(col, h, tin, tout, cbc);
    rs_array(tout, h, -sh[1]);
    for (int r = 0; r < h; r++) buf[r * w + (lr ? (w - 1 - c) : c)] = tout[r];
  }
  for (int r = 0; r < h; r++) {
    or_fwd_txfm1d(row, w, buf + r * w, tout, cbr);
    rs_array(tout, w, -sh[2]);
    for (int c = 0; c < w; c++)
      out[r * w + c] = rect ? rshift64((int64_t)tout[c] * 5793, 12) : tout[c];
  }
  free(buf);
  return 1;
}

/* dsp/inv_txfm.rs:1888 inv_txfm_shift */
static void inv_shift(int w, int h, int s[2]) {
  int lw = ilog2(w), lh = ilog2(h);
  int lo = lw < lh ? lw : lh, hi = lw < lh ? lh : lw;
  s[1] = -4;
  if (lo == hi) s[0] = (lo == 2) ? 0 : (lo == 3 ? -1 : -2);
  else if (hi - lo == 1) s[0] = (lo == 2) ? 0 : -1;
  else s[0] = (lo == 2) ? -1 : -2;
}

/* dsp/inv_txfm.rs:56 highbd_wraplow */
static int32_t wraplow(int32_t v, int bd) {
  int32_t mx = (1 << (7 + bd)) - 1 + (914 << (bd - 7));
  return v < -mx - 1 ? -mx - 1 : (v > mx ? mx : v);
}

/* dsp/inv_txfm.rs:1816 inv_txfm2d_core + txfm_dispatch.rs:55 (64-dim remap). */
int or_inv_txfm2d(const int32_t *in, int32_t *out, int stride, int tx_size, int tx_type, int bd) {
  int w, h, col, row, ud, lr, sh[2];
  if (tx_size < 0 || tx_size >= OR_TX_SIZES_ALL || tx_type < 0 || tx_type > 15) return 0;
  if (bd != 8 && bd != 10) return 0;
  or_tx_size_dims(tx_size, &w, &h);
  tx_type_cfg(tx_type, &col, &row, &ud, &lr);
  if (!kind_ok(col, h) || !kind_ok(row, w)) return 0;
  init_cospi();
  inv_shift(w, h, sh);
  int row_rng = bd + 8, col_rng = (bd + 6) > 16 ? bd + 6 : 16;
  int rect = abs(ilog2(w) - ilog2(h)) == 1;
  int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * w * h);
  int32_t tin[64], tout[64];
  for (int r = 0; r < h; r++) {
    for (int c = 0; c < w; c++) {
      int32_t v = (r < 32 && c < 32) ? in[r * stride + c] : 0;
      tin[c] = rect ? rshift64((int64_t)v * 2896, 12) : v;
      tin[c] = clampv(tin[c], row_rng);
    }
    or_inv_txfm1d(row, w, tin, buf + r * w, row_rng);
    rs_array(buf + r * w, w, -sh[0]);
  }
  for (int c = 0; c < w; c++) {
    for (int r = 0; r < h; r++) tin[r] = clampv(buf[r * w + (lr ? (w - 1 - c) : c)], col_rng);
    or_inv_txfm1d(col, h, tin, tout, col_rng);
    rs_array(tout, h, -sh[1]);
    for (int r = 0; r < h; r++) out[r * stride + c] = wraplow(tout[ud ? (h - 1 - r) : r], bd);
  }
  free(buf);
  return 1;
}
