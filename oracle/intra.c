/*
 * Oracle: AV1 intra predictors as zenav1-svt computes them, for 8- and
 * 10-bit samples (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h).
 *
 *   neighbour fill (non-directional / filter intra)
 *       enc/partition.rs:250 extract_neighbors_tiled (+ _hbd fallbacks)
 *   DC / V / H / smooth / smooth_v / smooth_h / paeth
 *       dsp/intra_pred.rs:18-154; hbd twins dsp/hbd.rs:169-361
 *       (note: the 10-bit Paeth breaks ties left-first, dsp/hbd.rs:203,
 *        the 8-bit one top-first, dsp/intra_pred.rs:282)
 *   directional: edge assembly, corner/edge filter, upsampling, z1/z2/z3
 *       enc/intra_edge.rs:910 dr_predict; dsp/intra_pred.rs:545-866;
 *       hbd dsp/hbd.rs:436-700
 *   filter intra  dsp/intra_pred.rs:988 (hbd dsp/hbd.rs:784)
 *   CfL           dsp/intra_pred.rs:1088,1110,1137 (hbd dsp/hbd.rs:855,889)
 *   routing       lf/predict.rs:49 predict_unit
 */
#include <stdlib.h>
#include <string.h>

#include "av1_intra_tables.h"
#include "zav1_oracle.h"

#define EDGE_ORIGIN 16
#define EDGE_LEN 160

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static const uint8_t *sm_weights(int n) {
  switch (n) {
    case 4: return SM_WEIGHTS_ALL + 0;
    case 8: return SM_WEIGHTS_ALL + 4;
    case 16: return SM_WEIGHTS_ALL + 12;
    case 32: return SM_WEIGHTS_ALL + 28;
    default: return SM_WEIGHTS_ALL + 60;
  }
}

/* dsp/intra_pred.rs:337/:347 get_dx / get_dy */
static int get_dx(int a) {
  if (a > 0 && a < 90) return DR_INTRA_DERIVATIVE[a];
  if (a > 90 && a < 180) return DR_INTRA_DERIVATIVE[180 - a];
  return 1;
}
static int get_dy(int a) {
  if (a > 90 && a < 180) return DR_INTRA_DERIVATIVE[a - 90];
  if (a > 180 && a < 270) return DR_INTRA_DERIVATIVE[270 - a];
  return 1;
}

/* dsp/intra_pred.rs:545 intra_edge_filter_strength */
int or_edge_filter_strength(int bs0, int bs1, int delta, int filt_type) {
  int d = delta < 0 ? -delta : delta, wh = bs0 + bs1, s = 0;
  if (filt_type == 0) {
    if (wh <= 8) { if (d >= 56) s = 1; }
    else if (wh <= 12) { if (d >= 40) s = 1; }
    else if (wh <= 16) { if (d >= 40) s = 1; }
    else if (wh <= 24) { if (d >= 8) s = 1; if (d >= 16) s = 2; if (d >= 32) s = 3; }
    else if (wh <= 32) { if (d >= 1) s = 1; if (d >= 4) s = 2; if (d >= 32) s = 3; }
    else if (d >= 1) s = 3;
  } else {
    if (wh <= 8) { if (d >= 40) s = 1; if (d >= 64) s = 2; }
    else if (wh <= 16) { if (d >= 20) s = 1; if (d >= 48) s = 2; }
    else if (wh <= 24) { if (d >= 4) s = 3; }
    else if (d >= 1) s = 3;
  }
  return s;
}

/* dsp/intra_pred.rs:610 use_intra_edge_upsample */
int or_use_upsample(int bs0, int bs1, int delta, int filt_type) {
  int d = delta < 0 ? -delta : delta, wh = bs0 + bs1;
  if (d <= 0 || d >= 40) return 0;
  return filt_type ? (wh <= 8) : (wh <= 16);
}

/* dsp/intra_pred.rs:626 filter_intra_edge on p[start .. start+sz) */
static void filter_edge(int32_t *p, int start, int sz, int strength) {
  static const int K[3][5] = {{0, 4, 8, 4, 0}, {0, 5, 6, 5, 0}, {2, 4, 4, 4, 2}};
  if (strength == 0) return;
  int32_t e[EDGE_LEN];
  memcpy(e, p + start, sizeof(int32_t) * sz);
  for (int i = 1; i < sz; i++) {
    int32_t s = 0;
    for (int j = 0; j < 5; j++) s += e[clampi(i - 2 + j, 0, sz - 1)] * K[strength - 1][j];
    p[start + i] = (s + 8) >> 4;
  }
}

/* dsp/intra_pred.rs:657 upsample_intra_edge (writes p[o-2 .. o+2sz-2]) */
static void upsample_edge(int32_t *p, int o, int sz, int maxv) {
  int32_t in[16 + 3];
  in[0] = p[o - 1];
  in[1] = p[o - 1];
  for (int i = 0; i < sz; i++) in[2 + i] = p[o + i];
  in[sz + 2] = p[o + sz - 1];
  p[o - 2] = in[0];
  for (int i = 0; i < sz; i++) {
    int32_t s = -in[i] + 9 * in[i + 1] + 9 * in[i + 2] - in[i + 3];
    p[o + 2 * i - 1] = clampi((s + 8) >> 4, 0, maxv);
    p[o + 2 * i] = in[i + 2];
  }
}

static void dr_z1(int32_t *dst, int bw, int bh, const int32_t *above, int up, int dx, int maxv) {
  const int o = EDGE_ORIGIN;
  int max_base_x = ((bw + bh) - 1) << up, frac = 6 - up, inc = 1 << up, x = dx;
  for (int r = 0; r < bh; r++) {
    int base = x >> frac, shift = ((x << up) & 0x3F) >> 1;
    if (base >= max_base_x) {
      for (int rr = r; rr < bh; rr++)
        for (int c = 0; c < bw; c++) dst[rr * bw + c] = above[o + max_base_x];
      return;
    }
    for (int c = 0; c < bw; c++, base += inc) {
      if (base < max_base_x) {
        int32_t v = above[o + base] * (32 - shift) + above[o + base + 1] * shift;
        dst[r * bw + c] = clampi((v + 16) >> 5, 0, maxv);
      } else {
        dst[r * bw + c] = above[o + max_base_x];
      }
    }
    x += dx;
  }
}

static void dr_z2(int32_t *dst, int bw, int bh, const int32_t *above, const int32_t *left, int up_a,
                  int up_l, int dx, int dy, int maxv) {
  const int o = EDGE_ORIGIN;
  int min_base_x = -(1 << up_a), fx = 6 - up_a, fy = 6 - up_l;
  for (int r = 0; r < bh; r++)
    for (int c = 0; c < bw; c++) {
      int y = r + 1, x = (c << 6) - y * dx, base = x >> fx;
      int32_t v;
      if (base >= min_base_x) {
        int shift = ((x * (1 << up_a)) & 0x3F) >> 1;
        v = above[o + base] * (32 - shift) + above[o + base + 1] * shift;
      } else {
        int x2 = c + 1, y2 = (r << 6) - x2 * dy, base2 = y2 >> fy;
        int shift = ((y2 * (1 << up_l)) & 0x3F) >> 1;
        v = left[o + base2] * (32 - shift) + left[o + base2 + 1] * shift;
      }
      dst[r * bw + c] = clampi((v + 16) >> 5, 0, maxv);
    }
}

static void dr_z3(int32_t *dst, int bw, int bh, const int32_t *left, int up, int dy, int maxv) {
  const int o = EDGE_ORIGIN;
  int max_base_y = (bw + bh - 1) << up, frac = 6 - up, inc = 1 << up, y = dy;
  for (int c = 0; c < bw; c++) {
    int base = y >> frac, shift = ((y << up) & 0x3F) >> 1;
    for (int r = 0; r < bh; r++, base += inc) {
      if (base < max_base_y) {
        int32_t v = left[o + base] * (32 - shift) + left[o + base + 1] * shift;
        dst[r * bw + c] = clampi((v + 16) >> 5, 0, maxv);
      } else {
        for (; r < bh; r++) dst[r * bw + c] = left[o + max_base_y];
        break;
      }
    }
    y += dy;
  }
}

/* dsp/intra_pred.rs:810 dr_predictor_edged: final edge buffers with the
 * sample 0 of the row / column at EDGE_ORIGIN; angle in (0, 270). */
void or_dr_predict_edges(const int32_t *above, const int32_t *left, int w, int h, int angle, int up_a, int up_l,
                         int bd, int32_t *dst) {
  const int maxv = (1 << bd) - 1;
  if (angle > 0 && angle < 90) dr_z1(dst, w, h, above, up_a, get_dx(angle), maxv);
  else if (angle > 90 && angle < 180) dr_z2(dst, w, h, above, left, up_a, up_l, get_dx(angle), get_dy(angle), maxv);
  else if (angle > 180 && angle < 270) dr_z3(dst, w, h, left, up_l, get_dy(angle), maxv);
  else if (angle == 90)
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) dst[r * w + c] = above[EDGE_ORIGIN + c];
  else if (angle == 180)
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) dst[r * w + c] = left[EDGE_ORIGIN + r];
}

/* Plane sample fetch: 8-bit or 16-bit plane. */
static int32_t S(const void *p, int stride, int hbd, int x, int y) {
  return hbd ? ((const uint16_t *)p)[(long)y * stride + x] : ((const uint8_t *)p)[(long)y * stride + x];
}

/* enc/intra_edge.rs:910 dr_predict, sample-generic. */
static void predict_dr(const void *plane, int stride, int hbd, int bd, const or_intra_job *j, int p_angle,
                       int32_t *dst) {
  const int w = j->w, h = j->h, base = 128 << (bd - 8), maxv = (1 << bd) - 1;
  const int n_top = j->n_top_px, n_tr = j->n_topright_px, n_left = j->n_left_px, n_bl = j->n_bottomleft_px;
  int need_above, need_left;
  if (p_angle <= 90) { need_above = 1; need_left = 0; }
  else if (p_angle < 180) { need_above = 1; need_left = 1; }
  else { need_above = 0; need_left = 1; }
  const int need_right = p_angle < 90, need_bottom = p_angle > 180;
#define AREF(i) S(plane, stride, hbd, j->x + (i), j->y - 1)
#define LREF(i) S(plane, stride, hbd, j->x - 1, j->y + (i))
  if ((!need_above && n_left == 0) || (!need_left && n_top == 0)) {
    int32_t v = need_left ? (n_top > 0 ? AREF(0) : base + 1) : (n_left > 0 ? LREF(0) : base - 1);
    for (int i = 0; i < w * h; i++) dst[i] = v;
    return;
  }
  const int o = EDGE_ORIGIN;
  int32_t above[EDGE_LEN], left[EDGE_LEN];
  for (int i = 0; i < EDGE_LEN; i++) { above[i] = base - 1; left[i] = base + 1; }
  if (need_left) {
    int num_left = h + (need_bottom ? w : 0);
    if (n_left > 0) {
      int i;
      for (i = 0; i < n_left; i++) left[o + i] = LREF(i);
      if (need_bottom && n_bl > 0)
        for (; i < h + n_bl; i++) left[o + i] = LREF(i);
      for (int k = i; k < num_left; k++) left[o + k] = left[o + i - 1];
    } else if (n_top > 0) {
      for (int k = 0; k < num_left; k++) left[o + k] = AREF(0);
    }
  }
  if (need_above) {
    int num_top = w + (need_right ? h : 0);
    if (n_top > 0) {
      int i;
      for (i = 0; i < n_top; i++) above[o + i] = AREF(i);
      if (need_right && n_tr > 0)
        for (; i < w + n_tr; i++) above[o + i] = AREF(i);
      for (int k = i; k < num_top; k++) above[o + k] = above[o + i - 1];
    } else if (n_left > 0) {
      for (int k = 0; k < num_top; k++) above[o + k] = LREF(0);
    }
  }
  int32_t tl;
  if (n_top > 0 && n_left > 0) tl = S(plane, stride, hbd, j->x - 1, j->y - 1);
  else if (n_top > 0) tl = AREF(0);
  else if (n_left > 0) tl = LREF(0);
  else tl = base;
  above[o - 1] = tl;
  left[o - 1] = tl;
  int up_a = 0, up_l = 0;
  if (j->edge_filter) {
    if (p_angle != 90 && p_angle != 180) {
      if (need_above && need_left && (w + h >= 24)) {
        int32_t s = (left[o] * 5 + above[o - 1] * 6 + above[o] * 5 + 8) >> 4;
        above[o - 1] = s;
        left[o - 1] = s;
      }
      if (need_above && n_top > 0) {
        int st = or_edge_filter_strength(w, h, p_angle - 90, j->filt_type);
        filter_edge(above, o - 1, n_top + 1 + (need_right ? h : 0), st);
      }
      if (need_left && n_left > 0) {
        int st = or_edge_filter_strength(h, w, p_angle - 180, j->filt_type);
        filter_edge(left, o - 1, n_left + 1 + (need_bottom ? w : 0), st);
      }
    }
    up_a = or_use_upsample(w, h, p_angle - 90, j->filt_type);
    if (need_above && up_a) upsample_edge(above, o, w + (need_right ? h : 0), maxv);
    up_l = or_use_upsample(h, w, p_angle - 180, j->filt_type);
    if (need_left && up_l) upsample_edge(left, o, h + (need_bottom ? w : 0), maxv);
  }
  int dx = get_dx(p_angle), dy = get_dy(p_angle);
  if (p_angle > 0 && p_angle < 90) dr_z1(dst, w, h, above, up_a, dx, maxv);
  else if (p_angle > 90 && p_angle < 180) dr_z2(dst, w, h, above, left, up_a, up_l, dx, dy, maxv);
  else if (p_angle > 180 && p_angle < 270) dr_z3(dst, w, h, left, up_l, dy, maxv);
  else if (p_angle == 90) {
    for (int r = 0; r < h; r++) for (int c = 0; c < w; c++) dst[r * w + c] = above[o + c];
  } else {
    for (int r = 0; r < h; r++) for (int c = 0; c < w; c++) dst[r * w + c] = left[o + r];
  }
#undef AREF
#undef LREF
}

/* enc/partition.rs:250 extract_neighbors_tiled with bd-general fallbacks. */
static void neighbours(const void *plane, int stride, int hbd, int bd, const or_intra_job *j, int32_t *above,
                       int32_t *left, int32_t *top_left) {
  const int w = j->w, h = j->h, base = 128 << (bd - 8);
  const int has_above = j->n_top_px > 0, has_left = j->n_left_px > 0;
  int32_t a0 = has_above ? S(plane, stride, hbd, j->x, j->y - 1) : 0;
  int32_t l0 = has_left ? S(plane, stride, hbd, j->x - 1, j->y) : 0;
  if (has_above) {
    int32_t last = a0;
    for (int i = 0; i < w; i++) {
      if (i < j->n_top_px) last = S(plane, stride, hbd, j->x + i, j->y - 1);
      above[i] = last;
    }
  } else {
    for (int i = 0; i < w; i++) above[i] = has_left ? l0 : base - 1;
  }
  if (has_left) {
    int32_t last = l0;
    for (int i = 0; i < h; i++) {
      if (i < j->n_left_px) last = S(plane, stride, hbd, j->x - 1, j->y + i);
      left[i] = last;
    }
  } else {
    for (int i = 0; i < h; i++) left[i] = has_above ? a0 : base + 1;
  }
  if (has_above && has_left) *top_left = S(plane, stride, hbd, j->x - 1, j->y - 1);
  else if (has_above) *top_left = a0;
  else if (has_left) *top_left = l0;
  else *top_left = base;
}

static int32_t rpot_signed(int32_t v, int n) {
  return v < 0 ? -((-v + (1 << (n - 1))) >> n) : ((v + (1 << (n - 1))) >> n);
}

int or_intra_predict(const void *plane, int stride, int bd, const or_intra_job *j, int32_t *dst) {
  const int w = j->w, h = j->h, hbd = bd > 8, maxv = (1 << bd) - 1;
  if (w < 4 || w > 64 || h < 4 || h > 64 || j->mode > 12) return 0;
  if ((j->mode >= 3 && j->mode <= 8) || ((j->mode == 1 || j->mode == 2) && j->delta != 0)) {
    predict_dr(plane, stride, hbd, bd, j, MODE_TO_ANGLE[j->mode] + 3 * j->delta, dst);
    return 1;
  }
  int32_t above[64], left[64], tl;
  neighbours(plane, stride, hbd, bd, j, above, left, &tl);
  if (j->fi_mode < 5) {
    /* dsp/intra_pred.rs:988 predict_filter_intra on a 33x33 staging buffer */
    if (w > 32 || h > 32) return 0;
    int32_t buf[33][33];
    buf[0][0] = tl;
    for (int c = 0; c < w; c++) buf[0][c + 1] = above[c];
    for (int r = 0; r < h; r++) buf[r + 1][0] = left[r];
    const int8_t *taps = FILTER_INTRA_TAPS + j->fi_mode * 64;
    for (int r = 1; r < h + 1; r += 2)
      for (int c = 1; c < w + 1; c += 4) {
        int32_t p[7] = {buf[r - 1][c - 1], buf[r - 1][c], buf[r - 1][c + 1], buf[r - 1][c + 2],
                        buf[r - 1][c + 3], buf[r][c - 1], buf[r + 1][c - 1]};
        for (int k = 0; k < 8; k++) {
          int32_t v = 0;
          for (int t = 0; t < 7; t++) v += taps[k * 8 + t] * p[t];
          buf[r + (k >> 2)][c + (k & 3)] = clampi(rpot_signed(v, 4), 0, maxv);
        }
      }
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) dst[r * w + c] = buf[r + 1][c + 1];
    return 1;
  }
  const int has_above = j->n_top_px > 0, has_left = j->n_left_px > 0;
  switch (j->mode) {
    case 0: {
      uint32_t sum = 0, dc;
      if (has_above && has_left) {
        for (int i = 0; i < w; i++) sum += above[i];
        for (int i = 0; i < h; i++) sum += left[i];
        dc = (sum + (w + h) / 2) / (w + h);
      } else if (has_above) {
        for (int i = 0; i < w; i++) sum += above[i];
        dc = (sum + w / 2) / w;
      } else if (has_left) {
        for (int i = 0; i < h; i++) sum += left[i];
        dc = (sum + h / 2) / h;
      } else {
        dc = 128u << (bd - 8);
      }
      for (int i = 0; i < w * h; i++) dst[i] = (int32_t)dc;
      break;
    }
    case 1:
      for (int r = 0; r < h; r++) for (int c = 0; c < w; c++) dst[r * w + c] = above[c];
      break;
    case 2:
      for (int r = 0; r < h; r++) for (int c = 0; c < w; c++) dst[r * w + c] = left[r];
      break;
    case 9: {
      const uint8_t *wh = sm_weights(h), *ww = sm_weights(w);
      int32_t below = left[h - 1], right = above[w - 1];
      for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
          uint32_t p = (uint32_t)(wh[r] * above[c] + (256 - wh[r]) * below + ww[c] * left[r] +
                                  (256 - ww[c]) * right + 256) / 512;
          dst[r * w + c] = p > (uint32_t)maxv ? maxv : (int32_t)p;
        }
      break;
    }
    case 10: {
      const uint8_t *wh = sm_weights(h);
      int32_t below = left[h - 1];
      for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
          uint32_t p = (uint32_t)(wh[r] * above[c] + (256 - wh[r]) * below + 128) / 256;
          dst[r * w + c] = p > (uint32_t)maxv ? maxv : (int32_t)p;
        }
      break;
    }
    case 11: {
      const uint8_t *ww = sm_weights(w);
      int32_t right = above[w - 1];
      for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
          uint32_t p = (uint32_t)(ww[c] * left[r] + (256 - ww[c]) * right + 128) / 256;
          dst[r * w + c] = p > (uint32_t)maxv ? maxv : (int32_t)p;
        }
      break;
    }
    case 12:
      for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
          int32_t top = above[c], lf = left[r], base = top + lf - tl;
          int32_t pt = abs(base - top), pl = abs(base - lf), ptl = abs(base - tl), v;
          if (!hbd) v = (pt <= pl && pt <= ptl) ? top : (pl <= ptl ? lf : tl);
          else v = (pl <= pt && pl <= ptl) ? lf : (pt <= ptl ? top : tl);
          dst[r * w + c] = v;
        }
      break;
    default:
      return 0;
  }
  return 1;
}

/* CfL: dsp/intra_pred.rs:1088 subsample (Q3), :1110 subtract average,
 * :1137 predict; hbd dsp/hbd.rs:855,889.  luma: 2cw x 2ch reconstructed
 * luma; dc: cw x ch chroma DC prediction; alpha_q3 signed. */
void or_cfl_predict(const void *luma, int luma_stride, int bd, int cw, int ch, const int32_t *dc, int alpha_q3,
                    int32_t *dst) {
  const int hbd = bd > 8, maxv = (1 << bd) - 1;
  int16_t q3[32 * 32];
  for (int j = 0; j < ch; j++)
    for (int i = 0; i < cw; i++) {
      int32_t s = S(luma, luma_stride, hbd, 2 * i, 2 * j) + S(luma, luma_stride, hbd, 2 * i + 1, 2 * j) +
                  S(luma, luma_stride, hbd, 2 * i, 2 * j + 1) + S(luma, luma_stride, hbd, 2 * i + 1, 2 * j + 1);
      q3[j * 32 + i] = (int16_t)(s * 2);
    }
  int32_t sum = 0;
  for (int j = 0; j < ch; j++)
    for (int i = 0; i < cw; i++) sum += q3[j * 32 + i];
  int l2 = 0;
  while ((1 << l2) < cw * ch) l2++;
  int32_t avg = (sum + ((1 << l2) >> 1)) >> l2;
  for (int j = 0; j < ch; j++)
    for (int i = 0; i < cw; i++) {
      int16_t ac = (int16_t)(q3[j * 32 + i] - avg);
      int32_t sc = rpot_signed(alpha_q3 * ac, 6);
      dst[j * cw + i] = clampi(sc + dc[j * cw + i], 0, maxv);
    }
}
