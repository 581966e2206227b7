/*
 * Oracle: the RDOQ trellis (svt_av1_optimize_b) on the path without
 * quantization matrices (TEST INFRASTRUCTURE ONLY — see zav1_oracle.h).
 *
 * Restates enc/quant.rs:444-1540 (rate helpers :636-:906, update_coeff_general
 * :966, update_coeff_simple :1056, update_coeff_eob :1120, update_skip :1281,
 * optimize_b :1312) and the coefficient contexts of
 * rust/crates/svtav1-entropy/src/coeff_c.rs (levels map :236-:298, nz_mag :306,
 * nz_map_ctx_offset_2d :344, nz_map_ctx_from_stats :380, nz_map_ctx :406,
 * br_ctx_eob :456, br_ctx :537, eob_pos_token :606), and the coefficient
 * rate cost_coeffs_txb (lf/coeff_rate.rs:60).  The cost tables are inputs
 * (built by the caller from its CDFs, enc/quant.rs:514-:632).
 */
#include <stdint.h>
#include <string.h>

#include "zav1_oracle.h"

/* coeff_c.rs:41-66 */
static const int TXW[19] = {4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64};
static const int TXH[19] = {4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16};
static const int SQR[19] = {0, 1, 2, 3, 4, 0, 0, 1, 1, 2, 2, 3, 3, 0, 0, 1, 1, 2, 2};
static const int SQR_UP[19] = {0, 1, 2, 3, 4, 1, 1, 2, 2, 3, 3, 4, 4, 2, 2, 3, 3, 4, 4};
static const int LOG2M4[19] = {0, 2, 4, 6, 6, 1, 1, 3, 3, 5, 5, 6, 6, 2, 2, 4, 4, 5, 5};

enum { CLS_2D = 0, CLS_HORIZ = 1, CLS_VERT = 2 };
enum { PAD_HOR = 4, PAD_TOP = 2, PAD_BOTTOM = 2, PAD_END = 16 };
#define LEVELS_LEN ((PAD_TOP + 32 + PAD_BOTTOM + 4) * (32 + PAD_HOR) + PAD_END)

/* coeff_c.rs:69 adjusted_tx_size */
int or_adjusted_tx_size(int t) {
  if (t == 4 || t == 11 || t == 12) return 3;
  if (t == 18) return 10;
  if (t == 17) return 9;
  return t;
}
static int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}
/* coeff_c.rs:119 txsize_entropy_ctx */
int or_txs_ctx(int t) { return (SQR[t] + SQR_UP[t] + 1) >> 1; }
int or_eob_multi_size(int t) { return LOG2M4[t]; }

static int msb(uint32_t v) { return 31 - __builtin_clz(v); }
static int clip3(int v) { return v < 3 ? v : 3; }

/* coeff_c.rs:306 get_nz_mag; b = padded position of the coefficient */
static int nz_mag(const uint8_t *lv, int b, int bwl, int cls) {
  const int st = (1 << bwl) + PAD_HOR;
  int mag = clip3(lv[b + 1]) + clip3(lv[b + st]);
  if (cls == CLS_2D) {
    mag += clip3(lv[b + st + 1]) + clip3(lv[b + 2]) + clip3(lv[b + 2 * st]);
  } else if (cls == CLS_VERT) {
    mag += clip3(lv[b + 2 * st]) + clip3(lv[b + 3 * st]) + clip3(lv[b + 4 * st]);
  } else {
    mag += clip3(lv[b + 2]) + clip3(lv[b + 3]) + clip3(lv[b + 4]);
  }
  return mag;
}
/* coeff_c.rs:344 nz_map_ctx_offset_2d (shape from the unadjusted size) */
static int offset_2d(int tx_size, int ci) {
  if (ci == 0) return 0;
  const int bwl = log2i(TXW[or_adjusted_tx_size(tx_size)]);
  const int w = TXW[tx_size], h = TXH[tx_size];
  const int row = ci >> bwl, col = ci - (row << bwl);
  if (w < h) {
    if (row < 2) return 11;
  } else if (w > h && col < 2) {
    return 16;
  }
  if (row + col < 2) return 1;
  if (row + col < 4) return 6;
  return 21;
}
static int offset_1d(int pos) { return pos == 0 ? 26 : (pos == 1 ? 31 : 36); }

/* coeff_c.rs:406 nz_map_ctx (+ :380 from_stats); si = scan index */
static int lower_ctx(const uint8_t *lv, int ci, int bwl, int height, int si, int is_last, int tx_size, int cls) {
  if (is_last) {
    const int n = height << bwl;
    if (si == 0) return 0;
    if (si <= n / 8) return 1;
    if (si <= n / 4) return 2;
    return 3;
  }
  const int b = PAD_TOP * ((1 << bwl) + PAD_HOR) + ci + ((ci >> bwl) << 2);
  const int stats = nz_mag(lv, b, bwl, cls);
  if ((cls | ci) == 0) return 0;
  int ctx = (stats + 1) >> 1;
  if (ctx > 4) ctx = 4;
  const int row = ci >> bwl, col = ci - (row << bwl);
  if (cls == CLS_2D) return ctx + offset_2d(tx_size, ci);
  if (cls == CLS_HORIZ) return ctx + offset_1d(col);
  return ctx + offset_1d(row);
}
/* coeff_c.rs:456 br_ctx_eob */
static int br_ctx_eob(int c, int bwl, int cls) {
  const int row = c >> bwl, col = c - (row << bwl);
  if (c == 0) return 0;
  if ((cls == CLS_2D && row < 2 && col < 2) || (cls == CLS_HORIZ && col == 0) || (cls == CLS_VERT && row == 0))
    return 7;
  return 14;
}
/* coeff_c.rs:537 br_ctx */
static int br_ctx(const uint8_t *lv, int c, int bwl, int cls) {
  const int row = c >> bwl, col = c - (row << bwl);
  const int st = (1 << bwl) + PAD_HOR;
  const int pos = PAD_TOP * st + row * st + col;
  int mag = lv[pos + 1] + lv[pos + st];
  if (cls == CLS_2D) {
    mag += lv[pos + st + 1];
    mag = (mag + 1) >> 1;
    if (mag > 6) mag = 6;
    if (c == 0) return mag;
    if (row < 2 && col < 2) return mag + 7;
  } else if (cls == CLS_HORIZ) {
    mag += lv[pos + 2];
    mag = (mag + 1) >> 1;
    if (mag > 6) mag = 6;
    if (c == 0) return mag;
    if (col == 0) return mag + 7;
  } else {
    mag += lv[pos + 2 * st];
    mag = (mag + 1) >> 1;
    if (mag > 6) mag = 6;
    if (c == 0) return mag;
    if (row == 0) return mag + 7;
  }
  return mag + 14;
}

/* enc/quant.rs:636 RDCOST */
static int64_t rdcost(int64_t rdmult, int64_t rate, int64_t dist) { return ((rate * rdmult + 256) >> 9) + (dist << 7); }
/* enc/quant.rs:641 */
static int64_t coeff_dist(int32_t t, int32_t dq, int shift) {
  const int64_t d = ((int64_t)t - dq) * ((int64_t)1 << shift);
  return d * d;
}
/* enc/quant.rs:648 get_golomb_cost */
static int golomb_cost(int abs_qc) {
  if (abs_qc >= 15) {
    const int r = abs_qc - 14;
    return 512 * (2 * (msb((uint32_t)r) + 1) - 1);
  }
  return 0;
}
/* enc/quant.rs:660 get_br_cost */
static int br_cost(int level, const int32_t *lps) {
  int br = level - 3;
  if (br > 12) br = 12;
  return lps[br] + golomb_cost(level);
}
/* enc/quant.rs:736 get_br_cost_with_diff; the two Golomb tables (:666,
 * :701) in closed form: bits 512 * (2 * (msb(r) + 1) - 1); diff 512 at
 * r = 1 and 1024 at r = 2^k, k >= 1. */
static int br_cost_with_diff(int level, const int32_t *lps, int *diff) {
  int br = level - 3;
  if (br > 12) br = 12;
  int gb = 0;
  if (level <= 15) *diff += lps[br + 13];
  if (level >= 15) {
    const int r = level - 14;
    gb = 512 * (2 * (msb((uint32_t)r) + 1) - 1);
    *diff += r == 1 ? 512 : ((r & (r - 1)) == 0 ? 1024 : 0);
  }
  return lps[br] + gb;
}
/* enc/quant.rs:758 get_coeff_cost_general */
static int coeff_cost_general(int is_last, int ci, int abs_qc, int sign, int ctx, int dc_ctx,
                              const or_txb_costs *c, int bwl, int cls, const uint8_t *lv) {
  int cost = is_last ? c->base_eob_cost[ctx][(abs_qc < 3 ? abs_qc : 3) - 1] : c->base_cost[ctx][abs_qc < 3 ? abs_qc : 3];
  if (abs_qc) {
    cost += ci == 0 ? c->dc_sign_cost[dc_ctx][sign] : 512;
    if (abs_qc > 2) {
      const int b = is_last ? br_ctx_eob(ci, bwl, cls) : br_ctx(lv, ci, bwl, cls);
      cost += br_cost(abs_qc, c->lps_cost[b]);
    }
  }
  return cost;
}
/* enc/quant.rs:795 get_coeff_cost_eob */
static int coeff_cost_eob(int ci, int abs_qc, int sign, int ctx, int dc_ctx, const or_txb_costs *c, int bwl, int cls) {
  int cost = c->base_eob_cost[ctx][(abs_qc < 3 ? abs_qc : 3) - 1];
  if (abs_qc) {
    cost += ci == 0 ? c->dc_sign_cost[dc_ctx][sign] : 512;
    if (abs_qc > 2) cost += br_cost(abs_qc, c->lps_cost[br_ctx_eob(ci, bwl, cls)]);
  }
  return cost;
}
/* enc/quant.rs:822 get_two_coeff_cost_simple */
static int two_coeff_cost_simple(int ci, int abs_qc, int ctx, const or_txb_costs *c, int bwl, int cls,
                                 const uint8_t *lv, int *cost_low) {
  int cost = c->base_cost[ctx][abs_qc < 3 ? abs_qc : 3];
  int diff = abs_qc <= 3 ? c->base_cost[ctx][abs_qc + 4] : 0;
  if (abs_qc) {
    cost += 512;
    if (abs_qc > 2) {
      int bd = 0;
      cost += br_cost_with_diff(abs_qc, c->lps_cost[br_ctx(lv, ci, bwl, cls)], &bd);
      diff += bd;
    }
  }
  *cost_low = cost - diff;
  return cost;
}
/* coeff_c.rs:606 eob_pos_token */
static int eob_pos_token(int eob, int *extra) {
  static const uint8_t SMALL[33] = {0, 1, 2, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 5, 5,
                                    6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6};
  static const uint8_t LARGE[17] = {6, 7, 8, 8, 9, 9, 9, 9, 10, 10, 10, 10, 10, 10, 10, 10, 11};
  static const int16_t START[12] = {0, 1, 2, 3, 5, 9, 17, 33, 65, 129, 257, 513};
  int t;
  if (eob < 33) {
    t = SMALL[eob];
  } else {
    int e = (eob - 1) >> 5;
    if (e > 16) e = 16;
    t = LARGE[e];
  }
  *extra = eob - START[t];
  return t;
}
/* enc/quant.rs:851 get_eob_cost */
int or_rdoq_eob_cost(int eob, const or_eob_costs *e, const or_txb_costs *c, int cls) {
  int extra;
  const int pt = eob_pos_token(eob, &extra);
  int cost = e->eob_cost[cls != CLS_2D][pt - 1];
  if (pt > 2) {
    const int cnt = pt - 3, bit = (extra >> cnt) & 1;
    cost += c->eob_extra_cost[cnt][bit] + 512 * cnt;
  }
  return cost;
}

typedef struct {
  const or_txb_costs *c;
  const or_eob_costs *e;
  int64_t rdmult;
  int tx_size, cls, skip_ctx, dc_ctx, sharp, bwl, height, shift;
  int32_t dq[2];
  const uint16_t *scan;
  const int32_t *t;
  int32_t *q, *d;
  uint8_t lv[LEVELS_LEN];
} Rdoq;

static int lv_idx(const Rdoq *s, int ci) {
  return PAD_TOP * ((1 << s->bwl) + PAD_HOR) + ci + ((ci >> s->bwl) << 2);
}
static void qc_dqc_low(int abs_qc, int sign, int dqv, int shift, int32_t *qc, int32_t *dqc) {
  const int lo = abs_qc - 1;
  *qc = (-sign ^ lo) + sign;
  const int32_t dlo = (int32_t)(((int64_t)lo * dqv) >> shift);
  *dqc = (-sign ^ dlo) + sign;
}

/* enc/quant.rs:966 update_coeff_general */
static void update_general(Rdoq *s, int *accu_rate, int64_t *accu_dist, int si, int eob) {
  const int ci = s->scan[si], dqv = s->dq[ci != 0], qc = s->q[ci];
  const int is_last = si == eob - 1;
  const int ctx = lower_ctx(s->lv, ci, s->bwl, s->height, si, is_last, s->tx_size, s->cls);
  if (qc == 0) {
    *accu_rate += s->c->base_cost[ctx][0];
    return;
  }
  const int sign = qc < 0, abs_qc = qc < 0 ? -qc : qc;
  const int32_t tqc = s->t[ci], dqc = s->d[ci];
  const int64_t dist = coeff_dist(tqc, dqc, s->shift), dist0 = coeff_dist(tqc, 0, s->shift);
  const int rate = coeff_cost_general(is_last, ci, abs_qc, sign, ctx, s->dc_ctx, s->c, s->bwl, s->cls, s->lv);
  const int64_t rd = rdcost(s->rdmult, rate, dist);
  int32_t qc_low, dqc_low;
  int abs_low, rate_low;
  int64_t dist_low;
  if (abs_qc == 1) {
    abs_low = 0, qc_low = 0, dqc_low = 0, dist_low = dist0;
    rate_low = s->c->base_cost[ctx][0];
  } else {
    qc_dqc_low(abs_qc, sign, dqv, s->shift, &qc_low, &dqc_low);
    abs_low = abs_qc - 1;
    dist_low = coeff_dist(tqc, dqc_low, s->shift);
    rate_low = coeff_cost_general(is_last, ci, abs_low, sign, ctx, s->dc_ctx, s->c, s->bwl, s->cls, s->lv);
  }
  const int64_t rd_low = rdcost(s->rdmult, rate_low, dist_low);
  if (rd_low < rd) {
    s->q[ci] = qc_low;
    s->d[ci] = dqc_low;
    s->lv[lv_idx(s, ci)] = (uint8_t)(abs_low < 127 ? abs_low : 127);
    *accu_rate += rate_low;
    *accu_dist += dist_low - dist0;
  } else {
    *accu_rate += rate;
    *accu_dist += dist - dist0;
  }
}

/* enc/quant.rs:1056 update_coeff_simple */
static void update_simple(Rdoq *s, int *accu_rate, int si) {
  const int ci = s->scan[si], dqv = s->dq[ci != 0], qc = s->q[ci];
  const int ctx = lower_ctx(s->lv, ci, s->bwl, 0, si, 0, s->tx_size, s->cls);
  if (qc == 0) {
    *accu_rate += s->c->base_cost[ctx][0];
    return;
  }
  const int abs_qc = qc < 0 ? -qc : qc;
  const int32_t abs_tqc = s->t[ci] < 0 ? -s->t[ci] : s->t[ci];
  const int32_t abs_dqc = s->d[ci] < 0 ? -s->d[ci] : s->d[ci];
  int rate_low = 0;
  const int rate = two_coeff_cost_simple(ci, abs_qc, ctx, s->c, s->bwl, s->cls, s->lv, &rate_low);
  if (abs_dqc < abs_tqc) {
    *accu_rate += rate;
    return;
  }
  const int64_t dist = coeff_dist(abs_tqc, abs_dqc, s->shift);
  const int64_t rd = rdcost(s->rdmult, rate, dist);
  const int abs_low = abs_qc - 1;
  const int32_t abs_dqc_low = (int32_t)(((int64_t)abs_low * dqv) >> s->shift);
  const int64_t rd_low = rdcost(s->rdmult, rate_low, coeff_dist(abs_tqc, abs_dqc_low, s->shift));
  if (rd_low < rd) {
    const int sign = qc < 0;
    s->q[ci] = (-sign ^ abs_low) + sign;
    s->d[ci] = (-sign ^ abs_dqc_low) + sign;
    s->lv[lv_idx(s, ci)] = (uint8_t)(abs_low < 127 ? abs_low : 127);
    *accu_rate += rate_low;
  } else {
    *accu_rate += rate;
  }
}

/* enc/quant.rs:1120 update_coeff_eob */
static void update_eob(Rdoq *s, int *accu_rate, int64_t *accu_dist, int *eob, int *nz_num, int *nz_ci, int si) {
  const int ci = s->scan[si], dqv = s->dq[ci != 0], qc = s->q[ci];
  const int ctx = lower_ctx(s->lv, ci, s->bwl, 0, si, 0, s->tx_size, s->cls);
  if (qc == 0) {
    *accu_rate += s->c->base_cost[ctx][0];
    return;
  }
  int lower = 0;
  const int abs_qc = qc < 0 ? -qc : qc, sign = qc < 0;
  const int32_t tqc = s->t[ci], dqc = s->d[ci];
  const int64_t dist0 = coeff_dist(tqc, 0, s->shift);
  int64_t dist = coeff_dist(tqc, dqc, s->shift) - dist0;
  int rate = coeff_cost_general(0, ci, abs_qc, sign, ctx, s->dc_ctx, s->c, s->bwl, s->cls, s->lv);
  const int64_t rd = rdcost(s->rdmult, *accu_rate + rate, *accu_dist + dist);
  int32_t qc_low, dqc_low;
  int abs_low, rate_low;
  int64_t dist_low, rd_low;
  if (abs_qc == 1) {
    abs_low = 0, qc_low = 0, dqc_low = 0, dist_low = 0;
    rate_low = s->c->base_cost[ctx][0];
    rd_low = rdcost(s->rdmult, *accu_rate + rate_low, *accu_dist);
  } else {
    qc_dqc_low(abs_qc, sign, dqv, s->shift, &qc_low, &dqc_low);
    abs_low = abs_qc - 1;
    dist_low = coeff_dist(tqc, dqc_low, s->shift) - dist0;
    rate_low = coeff_cost_general(0, ci, abs_low, sign, ctx, s->dc_ctx, s->c, s->bwl, s->cls, s->lv);
    rd_low = rdcost(s->rdmult, *accu_rate + rate_low, *accu_dist + dist_low);
  }
  int lower_new_eob = 0;
  const int new_eob = si + 1;
  const int ctx_ne = lower_ctx(s->lv, ci, s->bwl, s->height, si, 1, s->tx_size, s->cls);
  const int ne_cost = or_rdoq_eob_cost(new_eob, s->e, s->c, s->cls);
  int rate_ce = ne_cost + coeff_cost_eob(ci, abs_qc, sign, ctx_ne, s->dc_ctx, s->c, s->bwl, s->cls);
  int64_t dist_ne = dist;
  int64_t rd_ne = rdcost(s->rdmult, rate_ce, dist_ne);
  if (abs_low > 0) {
    const int rate_ce_low = ne_cost + coeff_cost_eob(ci, abs_low, sign, ctx_ne, s->dc_ctx, s->c, s->bwl, s->cls);
    const int64_t rd_ne_low = rdcost(s->rdmult, rate_ce_low, dist_low);
    if (rd_ne_low < rd_ne) {
      lower_new_eob = 1;
      rd_ne = rd_ne_low;
      rate_ce = rate_ce_low;
      dist_ne = dist_low;
    }
  }
  if (rd_low < rd) {
    lower = 1;
    rate = rate_low;
    dist = dist_low;
  }
  if (!s->sharp && rd_ne < (rd_low < rd ? rd_low : rd)) {
    for (int k = 0; k < *nz_num; k++) {
      s->lv[lv_idx(s, nz_ci[k])] = 0;
      s->q[nz_ci[k]] = 0;
      s->d[nz_ci[k]] = 0;
    }
    *eob = new_eob;
    *nz_num = 0;
    *accu_rate = rate_ce;
    *accu_dist = dist_ne;
    lower = lower_new_eob;
  } else {
    *accu_rate += rate;
    *accu_dist += dist;
  }
  if (lower) {
    s->q[ci] = qc_low;
    s->d[ci] = dqc_low;
    s->lv[lv_idx(s, ci)] = (uint8_t)(abs_low < 127 ? abs_low : 127);
  }
  if (s->q[ci] != 0) nz_ci[(*nz_num)++] = ci;
}

/* enc/quant.rs:1312 optimize_b (iwt = None).  *eob > 0.  Returns the new eob. */
int or_optimize_b(const int32_t *tcoeff, int32_t *qcoeff, int32_t *dqcoeff, int eob, const uint16_t *scan,
                  const int32_t dequant[2], const or_txb_costs *c, const or_eob_costs *e, int64_t rdmult,
                  int tx_size, int tx_class, int txb_skip_ctx, int dc_sign_ctx, int sharpness_flag,
                  int cut_off_num, int cut_off_denum) {
  static __thread Rdoq S;
  Rdoq *s = &S;
  const int adj = or_adjusted_tx_size(tx_size);
  const int w = TXW[adj], h = TXH[adj];
  s->c = c, s->e = e, s->rdmult = rdmult, s->tx_size = tx_size, s->cls = tx_class;
  s->skip_ctx = txb_skip_ctx, s->dc_ctx = dc_sign_ctx, s->sharp = sharpness_flag;
  s->bwl = log2i(w), s->height = h, s->shift = or_tx_scale(tx_size);
  s->dq[0] = dequant[0], s->dq[1] = dequant[1];
  s->scan = scan, s->t = tcoeff, s->q = qcoeff, s->d = dqcoeff;
  /* txb_init_levels (coeff_c.rs:278): zero, then min(|q|, 127) at the origin */
  memset(s->lv, 0, sizeof(s->lv));
  const int st = w + PAD_HOR;
  for (int r = 0; r < h; r++)
    for (int col = 0; col < w; col++) {
      const int32_t v = qcoeff[r * w + col];
      const uint32_t a = v < 0 ? -(uint32_t)v : (uint32_t)v;
      s->lv[PAD_TOP * st + r * st + col] = (uint8_t)(a < 127 ? a : 127);
    }
  const int non_skip_cost = c->txb_skip_cost[txb_skip_ctx][0], skip_cost = c->txb_skip_cost[txb_skip_ctx][1];
  int accu_rate = or_rdoq_eob_cost(eob, e, c, tx_class);
  int64_t accu_dist = 0;
  int si = eob - 1;
  const int ci = scan[si], qc = qcoeff[ci];
  const int abs_qc = qc < 0 ? -qc : qc, sign = qc < 0;
  int nz_num = 1, nz_ci[5] = {ci, 0, 0, 0, 0};
  if (abs_qc >= 2) {
    update_general(s, &accu_rate, &accu_dist, si, eob);
    si--;
  } else {
    const int ctx = lower_ctx(s->lv, ci, s->bwl, h, si, 1, tx_size, tx_class);
    accu_rate += coeff_cost_eob(ci, abs_qc, sign, ctx, dc_sign_ctx, c, s->bwl, tx_class);
    const int64_t dist = coeff_dist(tcoeff[ci], dqcoeff[ci], s->shift), dist0 = coeff_dist(tcoeff[ci], 0, s->shift);
    accu_dist += dist - dist0;
    si--;
  }
  while (si >= 0 && nz_num <= 4) {
    update_eob(s, &accu_rate, &accu_dist, &eob, &nz_num, nz_ci, si);
    si--;
  }
  if (si == -1 && nz_num <= 4) {  /* update_skip (:1281) */
    const int64_t rd = rdcost(rdmult, accu_rate + non_skip_cost, accu_dist);
    const int64_t rd_ne = rdcost(rdmult, skip_cost, 0);
    if (!sharpness_flag && rd_ne < rd) {
      for (int k = 0; k < nz_num; k++) qcoeff[nz_ci[k]] = 0, dqcoeff[nz_ci[k]] = 0;
      accu_rate = 0;
      eob = 0;
    }
  }
  int si_end = 1;
  if (cut_off_num != 0) {
    uint32_t cut = (uint32_t)((w * h) >> 7);
    const uint32_t alt = (uint32_t)eob * (uint32_t)cut_off_num / (uint32_t)cut_off_denum;
    if (alt > cut) cut = alt;
    si_end = eob - (int)cut;
    if (si_end < 1) si_end = 1;
  }
  for (; si >= si_end; si--) update_simple(s, &accu_rate, si);
  if (si == 0) {
    int64_t dummy = 0;
    update_general(s, &accu_rate, &dummy, 0, eob);
  }
  return eob;
}

/* lf/coeff_rate.rs:60 cost_coeffs_txb at the full middle loop: the rate of
 * one quantized block (raster, coded size) with eob > 0; txt_rate is the
 * caller's tx-type rate (added for luma), the contexts those of the final
 * levels (coeff_c.rs:492 get_nz_map_contexts). */
int or_cost_coeffs_txb(const int32_t *qcoeff, int eob, const uint16_t *scan, const or_txb_costs *c,
                       const or_eob_costs *e, int tx_size, int tx_class, int plane_type, int txb_skip_ctx,
                       int dc_sign_ctx, int txt_rate) {
  return or_cost_coeffs_txb_fast(qcoeff, eob, scan, c, e, tx_size, tx_class, plane_type, txb_skip_ctx, dc_sign_ctx,
                                 txt_rate, 1);
}

/* The same with C's mds_fast_coeff_est_level (the partition pre-pass's
 * svt_aom_txb_estimate_coeff_bits_pd0 passes 2, enc/pd0.rs:976-1001): the
 * middle loop prices only scan indices min(eob - 2, eob / fast_level) down
 * to 1; fast_level 1 is the full loop. */
int or_cost_coeffs_txb_fast(const int32_t *qcoeff, int eob, const uint16_t *scan, const or_txb_costs *c,
                            const or_eob_costs *e, int tx_size, int tx_class, int plane_type, int txb_skip_ctx,
                            int dc_sign_ctx, int txt_rate, int fast_level) {
  const int adj = or_adjusted_tx_size(tx_size);
  const int w = TXW[adj], h = TXH[adj], bwl = log2i(w);
  static __thread uint8_t lv[LEVELS_LEN];
  memset(lv, 0, sizeof(lv));
  if (eob > 1) {
    const int st = w + PAD_HOR;
    for (int r = 0; r < h; r++)
      for (int col = 0; col < w; col++) {
        const int32_t v = qcoeff[r * w + col];
        const uint32_t a = v < 0 ? -(uint32_t)v : (uint32_t)v;
        lv[PAD_TOP * st + r * st + col] = (uint8_t)(a < 127 ? a : 127);
      }
  }
  int cost = c->txb_skip_cost[txb_skip_ctx][0];
  if (plane_type == 0) cost += txt_rate;
  cost += or_rdoq_eob_cost(eob, e, c, tx_class);
  /* one position's level cost (the closure of coeff_rate.rs:117) */
#define LEVEL_COST(pos, v, is_eob_pos, is_dc, ctx)                                              \
  do {                                                                                           \
    const int lvl = (v) < 0 ? -(v) : (v);                                                        \
    cost += (is_eob_pos) ? c->base_eob_cost[ctx][(lvl < 3 ? lvl : 3) - 1]                        \
                         : c->base_cost[ctx][lvl < 3 ? lvl : 3];                                 \
    if ((v) != 0) {                                                                              \
      cost += (is_dc) ? c->dc_sign_cost[dc_sign_ctx][(v) < 0] : 512;                             \
      if (lvl > 2) {                                                                             \
        const int b = br_ctx(lv, pos, bwl, tx_class);                                            \
        cost += c->lps_cost[b][lvl - 3 < 12 ? lvl - 3 : 12];                                     \
        if (lvl >= 15) cost += golomb_cost(lvl);                                                 \
      }                                                                                          \
    }                                                                                            \
  } while (0)
  if (eob == 1) {
    LEVEL_COST(0, qcoeff[0], 1, 1, 0);
    return cost;
  }
  {
    const int pos = scan[eob - 1];
    const int ctx = lower_ctx(lv, pos, bwl, h, eob - 1, 1, tx_size, tx_class);
    LEVEL_COST(pos, qcoeff[pos], 1, 0, ctx);
  }
  {
    const int ctx = lower_ctx(lv, 0, bwl, h, 0, 0, tx_size, tx_class);
    LEVEL_COST(0, qcoeff[0], 0, 1, ctx);
  }
#undef LEVEL_COST
  const int mid = eob / fast_level < eob - 2 ? eob / fast_level : eob - 2;
  for (int i = mid; i >= 1; i--) {
    const int pos = scan[i], v = qcoeff[pos], lvl = v < 0 ? -v : v;
    const int ctx = lower_ctx(lv, pos, bwl, h, i, 0, tx_size, tx_class);
    if (v != 0) cost += 512;
    if (lvl > 2) {
      const int b = br_ctx(lv, pos, bwl, tx_class);
      cost += c->base_cost[ctx][3];
      if (lvl - 3 < 12) cost += c->lps_cost[b][lvl - 3];
      else cost += golomb_cost(lvl) + c->lps_cost[b][12];
    } else {
      cost += c->base_cost[ctx][lvl];
    }
  }
  return cost;
}
