/* Partition pre-pass (light PD0) of the allintra presets: CPU restatement.
 * TEST INFRASTRUCTURE ONLY (see zav1_oracle.h).
 *
 * Reference: rust/crates/svtav1-encoder/src/pd0.rs (paths below relative to
 * rust/crates/svtav1-encoder/src/):
 *   pd0.rs:81    compute_b64_variance (85-entry map, even-row means)
 *   pd0.rs:144   blk_var_map
 *   pd0.rs:185   qp_th_scaling_factors
 *   pd0.rs:218   kf_full_lambda_8bit (the PD0 lambda at every bit depth)
 *   pd0.rs:355   max_block_size_allintra        pd0.rs:371 input_resolution_factor
 *   pd0.rs:441   pd0_detector_allintra_demotes  pd0.rs:456 lvl6_cost_allintra
 *   pd0.rs:697   tx_quant_core (DCT_DCT, 64-fold, quantize_b, frequency SSE)
 *   pd0.rs:1019  cost_coeffs_txb_pd0 (fast coefficient-rate level 2)
 *   pd0.rs:1075  check_is_subres_safe
 *   pd0.rs:1494-1700 the LVL_0 / LVL_1 / LVL_5 block costs
 *   pd0.rs:1702  Pd0Ctx::pick (svt_aom_pick_partition_pd0 +
 *                test_split_partition_pd0: parent first, early exits)
 *   pd0.rs:1890 / :1990 / :2144 the per-superblock entry points.
 * Scope: frames whose sides are multiples of 64 (complete 64x64 units; the
 * specification's forced edge splits and edge shapes do not arise).  The
 * plane is 8-bit: a 10-bit source enters MSB-truncated (v >> 2), as the
 * reference's partition pass reads it (pipeline.rs:1062).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "av1_part_cdfs.h"
#include "av1_qlookup_tables.h"
#include "zav1_oracle.h"

static const int LOG2_LUT[5] = {6, 5, 4, 3, 2};
static const int GRID_LUT[5] = {1, 2, 4, 8, 16};
static const int BASE_LUT[5] = {0, 1, 5, 21, 85};

static int ilog2i(int v) { return 31 - __builtin_clz((unsigned)v); }
static uint64_t div_round(uint64_t x, uint64_t y) { return (x + (y >> 1)) / y; }

/* pd0.rs:81 */
void or_b64_variance(const uint8_t *src, int stride, int ox, int oy, uint16_t v[85]) {
  uint64_t m8[64], s8[64], m16[16], s16[16], m32[4], s32[4];
  for (int by = 0; by < 8; by++)
    for (int bx = 0; bx < 8; bx++) {
      uint64_t sum = 0, sq = 0;
      for (int r = 0; r < 8; r += 2) {
        const uint8_t *p = src + (size_t)(oy + by * 8 + r) * stride + ox + bx * 8;
        for (int c = 0; c < 8; c++) {
          sum += p[c];
          sq += (uint64_t)p[c] * p[c];
        }
      }
      m8[by * 8 + bx] = sum << 3;
      s8[by * 8 + bx] = sq << 11;
    }
  for (int by = 0; by < 4; by++)
    for (int bx = 0; bx < 4; bx++) {
      const int f = by * 16 + bx * 2;
      m16[by * 4 + bx] = (m8[f] + m8[f + 1] + m8[f + 8] + m8[f + 9]) >> 2;
      s16[by * 4 + bx] = (s8[f] + s8[f + 1] + s8[f + 8] + s8[f + 9]) >> 2;
    }
  for (int by = 0; by < 2; by++)
    for (int bx = 0; bx < 2; bx++) {
      const int f = by * 8 + bx * 2;
      m32[by * 2 + bx] = (m16[f] + m16[f + 1] + m16[f + 4] + m16[f + 5]) >> 2;
      s32[by * 2 + bx] = (s16[f] + s16[f + 1] + s16[f + 4] + s16[f + 5]) >> 2;
    }
  const uint64_t m64 = (m32[0] + m32[1] + m32[2] + m32[3]) >> 2;
  const uint64_t s64 = (s32[0] + s32[1] + s32[2] + s32[3]) >> 2;
  v[0] = (uint16_t)((s64 - m64 * m64) >> 16);
  for (int i = 0; i < 4; i++) v[1 + i] = (uint16_t)((s32[i] - m32[i] * m32[i]) >> 16);
  for (int i = 0; i < 16; i++) v[5 + i] = (uint16_t)((s16[i] - m16[i] * m16[i]) >> 16);
  for (int i = 0; i < 64; i++) v[21 + i] = (uint16_t)((s8[i] - m8[i] * m8[i]) >> 16);
}

/* pd0.rs:144: map index of the square at (ox, oy) in its superblock and its
 * four quadrants (level 1..3 blocks) */
static int var_index(int sq, int ox, int oy, int sub[4]) {
  const int lvl = 6 - ilog2i(sq);
  const int idx = BASE_LUT[lvl] + (oy >> LOG2_LUT[lvl]) * GRID_LUT[lvl] + (ox >> LOG2_LUT[lvl]);
  if (sub && lvl + 1 < 4) {
    const int sl = lvl + 1, ss = LOG2_LUT[sl], sb = BASE_LUT[sl], sg = GRID_LUT[sl];
    const int sx = ox >> ss, sy = oy >> ss;
    sub[0] = sb + sy * sg + sx;
    sub[1] = sub[0] + 1;
    sub[2] = sb + (sy + 1) * sg + sx;
    sub[3] = sub[2] + 1;
  }
  return idx;
}

/* pd0.rs:185 */
void or_qp_th_scaling(int qp, uint32_t *qw, uint32_t *qwd) {
  if (qp >= 46) {
    const double ex = -((double)(qp > 40 ? qp : 40) - 35.0) / 10.0;
    *qw = (uint32_t)((1.05 - exp(ex)) * 10000.0);
    *qwd = 10000;
  } else {
    *qw = (uint32_t)(qp > 10 ? qp : 10);
    *qwd = 63;
  }
}

/* pd0.rs:224 / :249 (alt factors off, no delta-q, PSNR lambda weights) */
uint32_t or_kf_full_lambda_8bit(int qindex, int cli_qp) {
  const int64_t q = DC_QLOOKUP_8[qindex];
  int64_t rdmult = (int64_t)((3.3 + 0.0015 * (double)q) * (double)q * (double)q);
  rdmult = (rdmult * 150) >> 7;
  uint32_t lambda = (uint32_t)rdmult;
  const uint32_t w = cli_qp >= 56 ? 175 : (cli_qp >= 16 ? 150 : 0);
  if (w) lambda = (uint32_t)(((uint64_t)lambda * w) >> 7);
  return lambda;
}

/* pd0.rs:355 */
int or_max_block_size_allintra(int var64, int qp) {
  uint32_t qw, qwd;
  or_qp_th_scaling(qp, &qw, &qwd);
  const uint16_t th = (uint16_t)div_round(7500ull * qw, qwd);
  return var64 <= th ? 64 : 32;
}

/* pd0.rs:371 */
int or_input_resolution_factor(long pixels) {
  static const int F[7] = {0, 1, 2, 3, 4, 4, 4};
  int r;
  if (pixels < 0x28500) r = 0;
  else if (pixels < 0x4CE00) r = 1;
  else if (pixels < 0xA1400) r = 2;
  else if (pixels < 0x16DA00) r = 3;
  else if (pixels < 0x535200) r = 4;
  else if (pixels < 0x140A000) r = 5;
  else r = 6;
  return F[r];
}

/* pd0.rs:441 */
int or_pd0_detector_demotes(const uint16_t v[85], int qp) {
  const int32_t var64 = v[0];
  const int32_t var32 = (((int32_t)v[1] + v[2] + v[3] + v[4]) >> 2) * 4;
  int32_t s16 = 0;
  for (int i = 5; i < 21; i++) s16 += v[i];
  const int32_t var16 = (s16 >> 4) * 16;
  uint32_t qw, qwd;
  or_qp_th_scaling(qp, &qw, &qwd);
  const int32_t th = (int32_t)div_round(7500ull * qw, qwd);
  return abs(var32 - var64) < th && abs(var16 - var32) < th;
}

/* pd0.rs:456 */
uint64_t or_lvl6_cost(const uint16_t v[85], int sq, int ox, int oy, int qp) {
  uint32_t qw32, qwd32;
  or_qp_th_scaling(qp, &qw32, &qwd32);
  const uint64_t qw = qw32, qwd = qwd32;
  int sub[4];
  const int idx = var_index(sq, ox, oy, sub);
  const uint64_t bv = v[idx], area = (uint64_t)sq * sq;
  uint64_t bias = 1000;
  if (sq == 64) {
    const uint64_t t = bv / div_round(100 * qw, qwd);
    bias += 50 * (t < 10 ? t : 10);
  } else if (sq >= 16) {
    uint64_t mn = UINT64_MAX, mx = 0;
    for (int i = 0; i < 4; i++) {
      if (v[sub[i]] < mn) mn = v[sub[i]];
      if (v[sub[i]] > mx) mx = v[sub[i]];
    }
    const uint64_t t = bv / div_round(400 * qw, qwd), u = (mx - mn) / div_round(25 * qw, qwd);
    bias += 25 * (t < 10 ? t : 10);
    bias += 10 * (u < 10 ? u : 10);
  } else {
    const uint64_t t = bv / div_round(25 * qw, qwd);
    bias += 40 * (t < 10 ? t : 10);
  }
  return area * bias / 1000;
}

/* The frame-initial rates the pass prices its symbols with: PARTITION_SPLIT
 * per square size 8..64 at sub-context 0, PARTITION_NONE at context 0 (the
 * 8x8 row, used for every size: rd_cost.c:1344 in the reference's notes) and
 * skip = 0 at context 0 (pd0.rs:1305 build_m6_pd0_tables_from_ctx, :1460,
 * :1479, :1489 for the default tables). */
void or_pd0_default_rates(int32_t split_bits[4], int32_t *none_bits, int32_t *skip0_bits) {
  for (int s = 0; s < 4; s++) {
    int32_t cost[10] = {0};
    or_syntax_rate(cost, AV1_PARTITION_CDF[s * 4], s == 0 ? 4 : 10);
    split_bits[s] = cost[3];
    if (s == 0) *none_bits = cost[0];
  }
  int32_t sk[2] = {0};
  or_syntax_rate(sk, AV1_SKIP_CDF[0], 2);
  *skip0_bits = sk[0];
}

typedef struct {
  const uint8_t *src;
  int stride, sb_x, sb_y, tile_x0, tile_y0;
  const or_pd0_cfg *cfg;
  int mode; /* 0, 1, 5, 6 */
  int max_sq, min_sq;
  int subres; /* 255 undetermined, 0 off, 1 on */
  uint16_t var[85];
  const or_txb_costs *txb;   /* [10] */
  const or_eob_costs (*eob)[2];
  or_pd0_node *nodes;
} Pd0;

static int node_index(int sq, int ox, int oy) {
  const int lvl = 6 - ilog2i(sq);
  return BASE_LUT[lvl] + (oy / sq) * GRID_LUT[lvl] + ox / sq;
}

static int c_tx_size(int w, int h) {
  for (int t = 0; t < 19; t++) {
    int tw, th;
    or_tx_size_dims(t, &tw, &th);
    if (tw == w && th == h) return t;
  }
  return -1;
}

/* DC prediction of a w x h block at (x, y) of the plane from its source
 * neighbours (pd0_use_src_samples; partition.rs:213 extract_neighbors) */
static int32_t dc_from_source(const Pd0 *c, int x, int y, int w, int h) {
  int32_t above[64], left[64];
  const int ha = y > c->tile_y0, hl = x > c->tile_x0;
  for (int i = 0; i < w; i++) above[i] = ha ? c->src[(size_t)(y - 1) * c->stride + x + i] : 0;
  for (int i = 0; i < h; i++) left[i] = hl ? c->src[(size_t)(y + i) * c->stride + x - 1] : 0;
  return or_dc_value(above, left, w, h, ha, hl, 8);
}

/* pd0.rs:1075 */
static int subres_safe(const Pd0 *c, int x, int y, int32_t dc) {
  int64_t even = 0, odd = 0;
  for (int r = 0; r < 64; r++)
    for (int k = 0; k < 64; k++) {
      const int d = abs((int)c->src[(size_t)(y + r) * c->stride + x + k] - dc);
      if (r & 1) odd += d;
      else even += d;
    }
  const int64_t e1 = even > 1 ? even : 1, o1 = odd > 1 ? odd : 1;
  const int64_t dev = ((e1 - o1) * 100) / o1;
  return llabs(dev) <= 5;
}

/* pd0.rs:697 tx_quant_core + the rate and full cost of LVL_0 / LVL_1 /
 * LVL_5 (pd0.rs:1520-1690) */
static uint64_t encode_cost(Pd0 *c, int sq, int ox, int oy) {
  const or_pd0_cfg *cfg = c->cfg;
  const int x = c->sb_x + ox, y = c->sb_y + oy;
  const int32_t dc = dc_from_source(c, x, y, sq, sq);
  int step = 0;
  if (c->mode == 5) {
    if (sq == 64 && c->subres == 255) c->subres = (uint8_t)subres_safe(c, x, y, dc);
    step = c->subres == 1 ? 1 : 0;
  }
  const int th_ = sq >> step;
  const int tx = c_tx_size(sq, th_);
  int32_t *res = malloc(sizeof(int32_t) * sq * th_), *coef = malloc(sizeof(int32_t) * sq * th_);
  for (int r = 0; r < th_; r++)
    for (int k = 0; k < sq; k++) res[r * sq + k] = (int32_t)c->src[(size_t)(y + (r << step)) * c->stride + x + k] - dc;
  or_fwd_txfm2d(res, coef, sq, tx, 0);
  /* 64-point fold: keep the top-left 32 x min(h, 32), the rest's energy */
  uint64_t fold = 0;
  const int pw = sq < 32 ? sq : 32, ph = th_ < 32 ? th_ : 32;
  int32_t *pk = malloc(sizeof(int32_t) * pw * ph);
  for (int r = 0; r < th_; r++)
    for (int k = 0; k < sq; k++) {
      const int64_t v = coef[r * sq + k];
      if (r < ph && k < pw) pk[r * pw + k] = (int32_t)v;
      else fold += (uint64_t)(v * v);
    }
  const int qidx = c->mode == 1 ? cfg->qindex : (cfg->qindex + 8 > 255 ? 255 : cfg->qindex + 8);
  or_quant_row row;
  or_build_quant_row(qidx, 8, &row);
  const int ls = or_tx_scale(tx), n = pw * ph;
  int32_t *q = malloc(sizeof(int32_t) * n), *dq = malloc(sizeof(int32_t) * n);
  or_quantize_b(pk, n, &row, ls, 8, q, dq);
  uint16_t scan[1024];
  or_scan(tx, 0, scan);
  const int eob = or_eob(q, scan, n);
  uint64_t dist = or_sse_i32(pk, dq, n) + fold;
  const int shift = (1 - ls) * 2;
  dist = shift < 0 ? dist << (-shift) : dist >> shift;
  dist <<= step;
  uint64_t bits;
  if (c->mode == 1) {
    const int cw = sq < 32 ? sq : 32, ch = th_ < 32 ? th_ : 32, th = (cw * ch) >> 5;
    const int tctx = or_txs_ctx(tx);
    const or_txb_costs *tc = &c->txb[tctx * 2];
    if (cfg->coeff_rate_est_lvl >= 2 && eob < th) bits = 6000 + (uint64_t)eob * 500;
    else if (eob == 0) bits = (uint64_t)tc->txb_skip_cost[0][1];
    else
      bits = (uint64_t)or_cost_coeffs_txb_fast(q, eob, scan, tc, &c->eob[or_eob_multi_size(tx)][0], tx, 0, 0, 0, 0,
                                               or_intra_txt_rate(tx, 0, 0), 2);
  } else {
    bits = 5000 + (uint64_t)cfg->ires * 1600 + 100 * (uint64_t)eob;
  }
  free(res);
  free(coef);
  free(pk);
  free(q);
  free(dq);
  const uint64_t rate = bits + (uint64_t)cfg->skip0_bits + (uint64_t)cfg->none_bits;
  return or_rdcost(cfg->lambda, rate, dist);
}

uint64_t or_pd0_block_cost_sb(const uint8_t *src, int stride, int sb_x, int sb_y, const or_pd0_cfg *cfg, int mode,
                              int sq, int ox, int oy, int *subres_state) {
  Pd0 c;
  memset(&c, 0, sizeof(c));
  c.src = src, c.stride = stride, c.sb_x = sb_x, c.sb_y = sb_y, c.cfg = cfg, c.mode = mode;
  c.subres = (uint8_t)*subres_state;
  static __thread or_txb_costs txb[10];
  static __thread or_eob_costs eob[7][2];
  static __thread int built = -1;
  if (mode == 1 && built != cfg->qindex) {
    or_build_coeff_costs(cfg->qindex, txb, eob);
    built = cfg->qindex;
  }
  c.txb = txb;
  c.eob = (const or_eob_costs(*)[2])eob;
  or_b64_variance(src, stride, sb_x, sb_y, c.var);
  const uint64_t v = mode == 6 ? or_lvl6_cost(c.var, sq, ox, oy, cfg->qp) : encode_cost(&c, sq, ox, oy);
  *subres_state = c.subres;
  return v;
}

/* pd0.rs:1702 Pd0Ctx::pick for interior nodes (complete 64x64 units) */
static uint64_t pick(Pd0 *c, int sq, int ox, int oy) {
  const int ni = node_index(sq, ox, oy);
  or_pd0_node *nd = &c->nodes[ni];
  nd->flags |= OR_PD0_VISITED;
  const int tested = sq <= c->max_sq && sq >= c->min_sq;
  uint64_t pc = 0;
  if (tested) {
    pc = c->mode == 6 ? or_lvl6_cost(c->var, sq, ox, oy, c->cfg->qp) : encode_cost(c, sq, ox, oy);
    nd->flags |= OR_PD0_TESTED;
    nd->cost = pc;
  }
  if (sq <= c->min_sq) return pc;
  const or_pd0_cfg *cfg = c->cfg;
  const int slot = ilog2i(sq) - 3;
  uint64_t split = 0;
  if (c->mode == 1) split = or_rdcost(cfg->lambda, (uint64_t)cfg->split_bits[slot], 0);
  else if (c->mode != 6) split = or_rdcost(cfg->lambda, 2 * (uint64_t)cfg->split_bits[slot], 0);
  const int half = sq / 2;
  int valid = 1;
  for (int i = 0; i < 4; i++) {
    if (c->mode != 6 && tested) {
      const uint64_t th = i == 0 ? 50 : 1000;
      /* (pc * th * 1000 <= split * 10^6) */
      if ((unsigned __int128)pc * th * 1000 <= (unsigned __int128)split * 1000000) {
        valid = 0;
        break;
      }
    }
    split += pick(c, half, ox + (i & 1) * half, oy + (i >> 1) * half);
  }
  if (!valid) return pc;
  if (tested && (unsigned __int128)pc * 1000 <= (unsigned __int128)split * 1000) return pc;
  nd->flags |= OR_PD0_SPLIT;
  return split;
}

static void leaf_fill(const Pd0 *c, int sq, int ox, int oy, uint8_t *map) {
  const or_pd0_node *nd = &c->nodes[node_index(sq, ox, oy)];
  if (nd->flags & OR_PD0_SPLIT) {
    const int h = sq / 2;
    for (int i = 0; i < 4; i++) leaf_fill(c, h, ox + (i & 1) * h, oy + (i >> 1) * h, map);
    return;
  }
  const int l = ilog2i(sq);
  for (int r = oy / 4; r < (oy + sq) / 4; r++)
    for (int k = ox / 4; k < (ox + sq) / 4; k++) map[r * 16 + k] = (uint8_t)l;
}

int or_pd0_sb(const uint8_t *src, int stride, int sb_x, int sb_y, const or_pd0_cfg *cfg, or_pd0_node nodes[341],
              uint8_t leaf_map[256], int *mode_out) {
  Pd0 c;
  memset(&c, 0, sizeof(c));
  c.src = src, c.stride = stride, c.sb_x = sb_x, c.sb_y = sb_y, c.cfg = cfg;
  if (cfg->level == 1 && cfg->tile_w > 0) c.tile_x0 = sb_x - sb_x % cfg->tile_w;
  if (cfg->level == 1 && cfg->tile_h > 0) c.tile_y0 = sb_y - sb_y % cfg->tile_h;
  static __thread or_txb_costs txb[10];
  static __thread or_eob_costs eob[7][2];
  static __thread int built = -1;
  if (cfg->level == 1 && built != cfg->qindex) {
    or_build_coeff_costs(cfg->qindex, txb, eob);
    built = cfg->qindex;
  }
  c.txb = txb;
  c.eob = (const or_eob_costs(*)[2])eob;
  or_b64_variance(src, stride, sb_x, sb_y, c.var);
  c.nodes = nodes;
  memset(nodes, 0, sizeof(or_pd0_node) * 341);
  c.min_sq = cfg->min_sq;
  c.subres = 255;
  switch (cfg->level) {
    case 0: /* pd0.rs:1990 (bd10-forced full-RD level: subres off, no detector) */
      c.mode = 0;
      c.subres = 0;
      c.max_sq = or_max_block_size_allintra(c.var[0], cfg->qp);
      break;
    case 1: /* pd0.rs:2144 */
      c.mode = 1;
      c.max_sq = cfg->cap_max_block ? or_max_block_size_allintra(c.var[0], cfg->qp) : 64;
      break;
    case 5: /* pd0.rs:1890: the detector picks LVL_5 or LVL_6 */
      c.mode = or_pd0_detector_demotes(c.var, cfg->qp) ? 5 : 6;
      c.max_sq = or_max_block_size_allintra(c.var[0], cfg->qp);
      break;
    default: return -1;
  }
  if (cfg->max_tx < c.max_sq) c.max_sq = cfg->max_tx;
  pick(&c, 64, 0, 0);
  leaf_fill(&c, 64, 0, 0, leaf_map);
  if (mode_out) *mode_out = c.mode;
  return 0;
}

int or_pd0_frame(const uint8_t *src, int stride, int width, int height, const or_pd0_cfg *cfg, or_pd0_node *nodes,
                 uint8_t *leaf_map, uint8_t *modes) {
  if (width % 64 || height % 64) return -1;
  const int sw = width / 64, sh = height / 64;
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < sw * sh; s++) {
    int m = 0;
    or_pd0_sb(src, stride, (s % sw) * 64, (s / sw) * 64, cfg, nodes + (size_t)s * 341, leaf_map + (size_t)s * 256,
              &m);
    if (modes) modes[s] = (uint8_t)m;
  }
  return 0;
}
