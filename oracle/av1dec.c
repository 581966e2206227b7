/* Oracle (TEST INFRASTRUCTURE ONLY -- see zav1_oracle.h): an AV1 decoder
 * for the still-image subset the library's bitstream writer emits, written
 * from the AV1 specification's decoding process, so that the writer's
 * output is checked by decoding it: the parsed syntax against what was
 * written and the reconstruction against the encoder's.
 *
 *   OBUs                  spec 5.3 (obu_header, leb128 sizes)
 *   sequence header       spec 5.5 (reduced still-picture header, color
 *                         config with mono_chrome)
 *   frame header          spec 5.9.2 uncompressed_header for a reduced
 *                         still picture; tile_info 5.9.15; quantization,
 *                         segmentation, delta q, loop filter, tx mode
 *   symbol decoder        spec 8.2.2 init_symbol, 8.2.6 read_symbol with the
 *                         CDFs in the specification's increasing form and
 *                         its adaptation; read_bool
 *   tile decode           spec 5.11: clear contexts, superblocks, decode
 *                         partition (NONE / SPLIT), intra frame mode info
 *                         (skip, y mode, angle delta, filter intra),
 *                         residual (TX_MODE_LARGEST: one transform per
 *                         block), coeffs 5.11.39 with the all_zero, dc_sign
 *                         contexts of 8.3.2 and the coefficient contexts
 *   prediction            spec 7.11.2 availability from BlockDecoded (the
 *                         flags the decoder keeps per superblock), the
 *                         predictor itself the oracle's pinned or_intra_predict
 *   reconstruction        spec 7.12.3 dequantization (Dc/Ac_Qlookup, the
 *                         0xFFFFFF mask, dqDenom, clip), the oracle's pinned
 *                         inverse transform, Clip1 of prediction + residual
 *
 * Subset: monochrome, no segmentation / delta q / palette / intra block
 * copy / CDEF / restoration, TX_MODE_LARGEST, square partitions NONE and
 * SPLIT; anything else is reported as unsupported (-2). */
#include <stdlib.h>
#include <string.h>

#include "av1_coef_cdfs.h"
#include "av1_part_cdfs.h"
#include "av1_qlookup_tables.h"
#include "zav1_oracle.h"

/* ---- bit reader for headers (f(n)) ------------------------------------------ */
typedef struct {
  const uint8_t *b;
  size_t n, pos; /* pos in bits */
} BR;
static uint32_t f(BR *r, int n) {
  uint32_t v = 0;
  for (int i = 0; i < n; i++) {
    const size_t p = r->pos++;
    const int bit = p / 8 < r->n ? (r->b[p / 8] >> (7 - p % 8)) & 1 : 0;
    v = (v << 1) | (uint32_t)bit;
  }
  return v;
}
static void byte_align(BR *r) { r->pos = (r->pos + 7) & ~(size_t)7; }
static size_t leb128(const uint8_t *b, size_t n, size_t *i) {
  size_t v = 0;
  for (int k = 0; k < 8 && *i < n; k++) {
    const uint8_t x = b[(*i)++];
    v |= (size_t)(x & 0x7f) << (7 * k);
    if (!(x & 0x80)) break;
  }
  return v;
}

/* ---- symbol decoder (spec 8.2) ---------------------------------------------- */
typedef struct {
  const uint8_t *d;
  size_t sz, pos; /* bits consumed from d */
  uint32_t value, range;
  int max_bits;
} SD;
static uint32_t sd_bits(SD *s, int n) {
  uint32_t v = 0;
  for (int i = 0; i < n; i++) {
    const size_t p = s->pos++;
    v = (v << 1) | (uint32_t)(p / 8 < s->sz ? (s->d[p / 8] >> (7 - p % 8)) & 1 : 0);
  }
  return v;
}
static void init_symbol(SD *s, const uint8_t *d, size_t sz) {
  s->d = d, s->sz = sz, s->pos = 0;
  const int nb = sz * 8 < 15 ? (int)(sz * 8) : 15;
  const uint32_t buf = sd_bits(s, nb);
  const uint32_t padded = buf << (15 - nb);
  s->value = ((1u << 15) - 1) ^ padded;
  s->range = 1u << 15;
  s->max_bits = (int)(8 * sz) - 15;
}
static int floor_log2(uint32_t x) { return 31 - __builtin_clz(x); }
/* cdf: the specification's form, cdf[i] = 32768 * P(X <= i), cdf[n-1] =
 * 32768, the adaptation counter at cdf[n] */
static int read_symbol(SD *s, uint16_t *cdf, int n, int adapt) {
  uint32_t cur = s->range, prev;
  int sym = -1;
  do {
    sym++;
    prev = cur;
    const uint32_t fv = (1u << 15) - cdf[sym];
    cur = ((s->range >> 8) * (fv >> 6) >> 1) + 4u * (uint32_t)(n - sym - 1);
  } while (s->value < cur);
  s->range = prev - cur;
  s->value -= cur;
  const int bits = 15 - floor_log2(s->range);
  s->range <<= bits;
  const int nb = bits < (s->max_bits > 0 ? s->max_bits : 0) ? bits : (s->max_bits > 0 ? s->max_bits : 0);
  const uint32_t nd = sd_bits(s, nb);
  const uint32_t padded = nd << (bits - nb);
  s->value = padded ^ (((s->value + 1) << bits) - 1);
  s->max_bits -= bits;
  if (adapt) {
    const int rate = 3 + (cdf[n] > 15) + (cdf[n] > 31) + (floor_log2((uint32_t)n) < 2 ? floor_log2((uint32_t)n) : 2);
    uint32_t tmp = 0;
    for (int i = 0; i < n - 1; i++) {
      tmp = i == sym ? (1u << 15) : tmp;
      if (tmp < cdf[i]) cdf[i] -= (uint16_t)((cdf[i] - tmp) >> rate);
      else cdf[i] += (uint16_t)((tmp - cdf[i]) >> rate);
    }
    cdf[n] += cdf[n] < 32;
  }
  return sym;
}
static int read_bool(SD *s) {
  uint16_t c[3] = {1u << 14, 1u << 15, 0};
  return read_symbol(s, c, 2, 0);
}
/* default CDFs converted from the inverse form of the tables */
static void to_spec(uint16_t *dst, const uint16_t *icdf, int n) {
  for (int i = 0; i < n; i++) dst[i] = (uint16_t)(32768 - icdf[i]);
  dst[n] = 0;
}
typedef struct {
  uint16_t part[20][11], skip[3][3], kfy[5][5][14], angle[8][8], fi[22][3], fim[6], etx[3][4][13][17];
  uint16_t txb[5][13][3], eobx[5][2][9][3], dcs[2][3][3], base_eob[5][2][4][4], base[5][2][42][5], br[4][2][21][5];
  uint16_t eobp[7][2][2][12];
} Cdf;
static void cdf_init(Cdf *c, int qidx) {
  const int q = qidx <= 20 ? 0 : (qidx <= 60 ? 1 : (qidx <= 120 ? 2 : 3));
  for (int i = 0; i < 20; i++) to_spec(c->part[i], AV1_PARTITION_CDF[i], i < 4 ? 4 : (i >= 16 ? 8 : 10));
  for (int i = 0; i < 3; i++) to_spec(c->skip[i], AV1_SKIP_CDF[i], 2);
  for (int a = 0; a < 5; a++)
    for (int l = 0; l < 5; l++) to_spec(c->kfy[a][l], AV1_KF_Y_CDF[a][l], 13);
  for (int i = 0; i < 8; i++) to_spec(c->angle[i], AV1_ANGLE_DELTA_CDF[i], 7);
  for (int i = 0; i < 22; i++) to_spec(c->fi[i], AV1_FILTER_INTRA_CDF[i], 2);
  to_spec(c->fim, AV1_FILTER_INTRA_MODE_CDF, 5);
  for (int e = 1; e < 3; e++)
    for (int t = 0; t < 4; t++)
      for (int d = 0; d < 13; d++) to_spec(c->etx[e][t][d], AV1_INTRA_EXT_TX_CDF[e][t][d], e == 1 ? 7 : 5);
  for (int t = 0; t < 5; t++) {
    for (int k = 0; k < 13; k++) to_spec(c->txb[t][k], AV1_TXB_SKIP_CDF[q][t][k], 2);
    for (int p = 0; p < 2; p++) {
      for (int k = 0; k < 9; k++) to_spec(c->eobx[t][p][k], AV1_EOB_EXTRA_CDF[q][t][p][k], 2);
      for (int k = 0; k < 4; k++) to_spec(c->base_eob[t][p][k], AV1_COEFF_BASE_EOB_CDF[q][t][p][k], 3);
      for (int k = 0; k < 42; k++) to_spec(c->base[t][p][k], AV1_COEFF_BASE_CDF[q][t][p][k], 4);
      if (t < 4)
        for (int k = 0; k < 21; k++) to_spec(c->br[t][p][k], AV1_COEFF_BR_CDF[q][t][p][k], 4);
    }
  }
  for (int p = 0; p < 2; p++)
    for (int k = 0; k < 3; k++) to_spec(c->dcs[p][k], AV1_DC_SIGN_CDF[q][p][k], 2);
  for (int p = 0; p < 2; p++)
    for (int k = 0; k < 2; k++) {
      to_spec(c->eobp[0][p][k], AV1_EOB_FLAG_CDF16[q][p][k], 5);
      to_spec(c->eobp[1][p][k], AV1_EOB_FLAG_CDF32[q][p][k], 6);
      to_spec(c->eobp[2][p][k], AV1_EOB_FLAG_CDF64[q][p][k], 7);
      to_spec(c->eobp[3][p][k], AV1_EOB_FLAG_CDF128[q][p][k], 8);
      to_spec(c->eobp[4][p][k], AV1_EOB_FLAG_CDF256[q][p][k], 9);
      to_spec(c->eobp[5][p][k], AV1_EOB_FLAG_CDF512[q][p][k], 10);
      to_spec(c->eobp[6][p][k], AV1_EOB_FLAG_CDF1024[q][p][k], 11);
    }
}

/* ---- decoder state ------------------------------------------------------------ */
typedef struct {
  int W, H, bd, sb128, qidx, mi_cols, mi_rows;
  int tile_cols, tile_rows;
  int col_starts[65], row_starts[65];
  /* per mi */
  uint8_t *ymode, *skip, *wlog, *decoded_mi;
  void *frame;
  or_dec_block *blocks;
  int n_blocks, max_blocks;
  int32_t *levels;
  size_t n_levels, max_levels;
} Dec;

typedef struct {
  int mc0, mc1, mr0, mr1;
  uint8_t above_dc[16384], left_dc[16384];
  uint8_t bd_flags[34][34]; /* BlockDecoded of the current superblock, [y + 1][x + 1] */
  int sb_r, sb_c;
  Cdf cdf;
  SD sd;
} Tile;

static void st(Dec *d, int x, int y, int v) {
  if (d->bd > 8) ((uint16_t *)d->frame)[(size_t)y * d->W + x] = (uint16_t)v;
  else ((uint8_t *)d->frame)[(size_t)y * d->W + x] = (uint8_t)v;
}

static int tx_size_of(int size) { return size == 4 ? 0 : (size == 8 ? 1 : (size == 16 ? 2 : (size == 32 ? 3 : 4))); }
static const int MODE_CTX[13] = {0, 1, 2, 3, 4, 4, 4, 4, 3, 0, 1, 2, 0};
static const int FI_DIR[5] = {0, 1, 2, 6, 0};
/* Tx_Type_Intra_Inv_Set1 / Set2 */
static const int INV_SET1[7] = {9, 0, 10, 11, 3, 1, 2};
static const int INV_SET2[5] = {9, 0, 3, 1, 2};


/* spec 8.3.2 get_coeff_base_ctx (non-final positions) and get_br_ctx over the
 * levels decoded so far (n x n, raster; class 0 2D, 1 horizontal, 2
 * vertical) */
static const int SIG_REF[3][5][2] = {{{0, 1}, {1, 0}, {1, 1}, {0, 2}, {2, 0}},
                                     {{0, 1}, {1, 0}, {0, 2}, {0, 3}, {0, 4}},
                                     {{0, 1}, {1, 0}, {2, 0}, {3, 0}, {4, 0}}};
static const int MAG_REF[3][3][2] = {{{0, 1}, {1, 0}, {1, 1}}, {{0, 1}, {1, 0}, {0, 2}}, {{0, 1}, {1, 0}, {2, 0}}};
static const int BASE_CTX_OFFSET[5][5] = {
    {0, 1, 6, 6, 21}, {1, 6, 6, 21, 21}, {6, 6, 21, 21, 21}, {6, 21, 21, 21, 21}, {21, 21, 21, 21, 21}};
static int coeff_base_ctx(const uint8_t *lv, int n, int pos, int cls) {
  const int row = pos / n, col = pos % n;
  int mag = 0;
  for (int k = 0; k < 5; k++) {
    const int rr = row + SIG_REF[cls][k][0], cc = col + SIG_REF[cls][k][1];
    if (rr < n && cc < n) mag += lv[rr * n + cc] < 3 ? lv[rr * n + cc] : 3;
  }
  const int ctx = (mag + 1) >> 1 < 4 ? (mag + 1) >> 1 : 4;
  if (cls == 0) {
    if (row == 0 && col == 0) return 0;
    return ctx + BASE_CTX_OFFSET[row < 4 ? row : 4][col < 4 ? col : 4];
  }
  const int idx = cls == 2 ? row : col;
  return ctx + 26 + 5 * (idx < 2 ? idx : 2);
}
static int coeff_br_ctx(const uint8_t *lv, int n, int pos, int cls) {
  const int row = pos / n, col = pos % n;
  int mag = 0;
  for (int k = 0; k < 3; k++) {
    const int rr = row + MAG_REF[cls][k][0], cc = col + MAG_REF[cls][k][1];
    if (rr < n && cc < n) mag += lv[rr * n + cc] < 15 ? lv[rr * n + cc] : 15;
  }
  mag = (mag + 1) >> 1 < 6 ? (mag + 1) >> 1 : 6;
  if (pos == 0) return mag;
  if (cls == 0) return (row < 2 && col < 2) ? mag + 7 : mag + 14;
  if (cls == 1) return col == 0 ? mag + 7 : mag + 14;
  return row == 0 ? mag + 7 : mag + 14;
}

/* spec 5.11.39 coeffs for a square luma transform equal to its block; the
 * levels (packed n x n, n = min(size, 32)) into q; returns eob and the dc
 * category */
static int read_coeffs(Dec *d, Tile *t, int x4, int y4, int size, int intra_dir, int *tx_type, int32_t *q,
                       int *dcat) {
  const int txs = tx_size_of(size), txs_ctx = txs, n = size < 32 ? size : 32, N = n * n;
  /* all_zero: the transform equals the block -> context 0 */
  const int all_zero = read_symbol(&t->sd, t->cdf.txb[txs_ctx][0], 2, 1);
  memset(q, 0, sizeof(int32_t) * N);
  *tx_type = 0;
  *dcat = 0;
  if (all_zero) return 0;
  /* transform_type: intra sets; DCT only from 32 points */
  const int set = size >= 32 ? 0 : (size == 16 ? 2 : 1);
  if (set > 0 && d->qidx > 0) {
    const int sym = read_symbol(&t->sd, t->cdf.etx[set][txs][intra_dir], set == 1 ? 7 : 5, 1);
    *tx_type = set == 1 ? INV_SET1[sym] : INV_SET2[sym];
  }
  const int cls = *tx_type < 10 ? 0 : ((*tx_type & 1) ? 1 : 2); /* 2D, horizontal, vertical */
  uint16_t scan[1024];
  or_scan(txs, *tx_type < 10 ? 0 : ((*tx_type & 1) ? 2 : 1), scan);
  const int ems = 2 * floor_log2((uint32_t)n) - 4;
  const int pt = read_symbol(&t->sd, t->cdf.eobp[ems][0][cls == 0 ? 0 : 1], 5 + ems, 1) + 1;
  int eob = pt < 2 ? pt : (1 << (pt - 2)) + 1;
  int shift = pt - 3;
  if (shift >= 0) {
    if (read_symbol(&t->sd, t->cdf.eobx[txs_ctx][0][pt - 3], 2, 1)) eob += 1 << shift;
    for (int i = 1; i < (pt - 2 > 0 ? pt - 2 : 0); i++) {
      shift = (pt - 2 > 0 ? pt - 2 : 0) - 1 - i;
      if (read_bool(&t->sd)) eob += 1 << shift;
    }
  }
  /* levels, from the last position (the coefficient contexts over the levels
   * decoded so far, held as the oracle's padded level map) */
  uint8_t lv[1024];
  memset(lv, 0, sizeof(lv));
  const int br_txs = txs_ctx < 3 ? txs_ctx : 3;
  for (int c = eob - 1; c >= 0; c--) {
    const int pos = scan[c];
    int level;
    if (c == eob - 1) {
      const int ctx = c == 0 ? 0 : (c <= N / 8 ? 1 : (c <= N / 4 ? 2 : 3));
      level = read_symbol(&t->sd, t->cdf.base_eob[txs_ctx][0][ctx], 3, 1) + 1;
    } else {
      level = read_symbol(&t->sd, t->cdf.base[txs_ctx][0][coeff_base_ctx(lv, n, pos, cls)], 4, 1);
    }
    if (level > 2) {
      const int bctx = coeff_br_ctx(lv, n, pos, cls);
      for (int idx = 0; idx < 4; idx++) {
        const int k = read_symbol(&t->sd, t->cdf.br[br_txs][0][bctx], 4, 1);
        level += k;
        if (k < 3) break;
      }
    }
    q[pos] = level;
    lv[pos] = (uint8_t)(level < 127 ? level : 127);
  }
  int cul = 0;
  for (int c = 0; c < eob; c++) {
    const int pos = scan[c];
    int sign = 0;
    if (q[pos] != 0) {
      if (c == 0) {
        /* dc_sign context: the above / left dc categories over the block */
        int s = 0;
        for (int k = 0; k < size / 4; k++) {
          const int a = t->above_dc[x4 + k], l = t->left_dc[y4 + k];
          s += (a == 1 ? -1 : (a == 2 ? 1 : 0)) + (l == 1 ? -1 : (l == 2 ? 1 : 0));
        }
        sign = read_symbol(&t->sd, t->cdf.dcs[0][s < 0 ? 1 : (s > 0 ? 2 : 0)], 2, 1);
      } else {
        sign = read_bool(&t->sd);
      }
    }
    if (q[pos] > 14) {
      int length = 0, bit;
      do {
        length++;
        bit = read_bool(&t->sd);
        if (length > 32) return -1;
      } while (!bit);
      uint32_t x = 1;
      for (int i = length - 2; i >= 0; i--) x = (x << 1) | (uint32_t)read_bool(&t->sd);
      q[pos] = (int32_t)x + 14;
    }
    if (pos == 0 && q[pos] > 0) *dcat = sign ? 1 : 2;
    q[pos] &= 0xFFFFF;
    cul += q[pos];
    if (sign) q[pos] = -q[pos];
  }
  (void)cul;
  return eob;
}

static int push_block(Dec *d, const or_dec_block *b, const int32_t *q, int nq) {
  if (d->n_blocks == d->max_blocks || d->n_levels + (size_t)nq > d->max_levels) return -1;
  or_dec_block *o = &d->blocks[d->n_blocks++];
  *o = *b;
  o->level_off = d->n_levels;
  memcpy(d->levels + d->n_levels, q, sizeof(int32_t) * nq);
  d->n_levels += (size_t)nq;
  return 0;
}

/* decode_block for an intra square block whose transform is the block */
static int decode_block(Dec *d, Tile *t, int r, int c, int size) {
  const int n4 = size / 4, W = d->mi_cols;
  const int up = r > t->mr0, left = c > t->mc0;
  const int sctx = (up ? d->skip[(r - 1) * W + c] : 0) + (left ? d->skip[r * W + c - 1] : 0);
  const int skip = read_symbol(&t->sd, t->cdf.skip[sctx], 2, 1);
  const int am = up ? d->ymode[(r - 1) * W + c] : 0, lm = left ? d->ymode[r * W + c - 1] : 0;
  const int mode = read_symbol(&t->sd, t->cdf.kfy[MODE_CTX[am]][MODE_CTX[lm]], 13, 1);
  int delta = 0, fi = 5;
  if (mode >= 1 && mode <= 8) delta = read_symbol(&t->sd, t->cdf.angle[mode - 1], 7, 1) - 3;
  if (mode == 0 && size <= 32) {
    const int bs = size == 8 ? 3 : (size == 16 ? 6 : 9);
    if (read_symbol(&t->sd, t->cdf.fi[bs], 2, 1)) fi = read_symbol(&t->sd, t->cdf.fim, 5, 1);
  }
  /* prediction, availability from BlockDecoded (spec 7.11.2) */
  const int x = c * 4, y = r * 4;
  const int sr = (r - t->sb_r), scl = (c - t->sb_c);
  const int have_ar = t->bd_flags[sr - 1 + 1][scl + n4 + 1];
  const int have_bl = t->bd_flags[sr + n4 + 1][scl - 1 + 1];
  const int maxX = d->mi_cols * 4 - 1, maxY = d->mi_rows * 4 - 1;
  or_intra_job jb;
  memset(&jb, 0, sizeof(jb));
  jb.x = x, jb.y = y, jb.w = (uint8_t)size, jb.h = (uint8_t)size, jb.mode = (uint8_t)mode, jb.delta = (int8_t)delta;
  jb.fi_mode = (uint8_t)fi, jb.edge_filter = 1;
  const int asm_ = up && (am >= 9 && am <= 11), lsm = left && (lm >= 9 && lm <= 11);
  jb.filt_type = (uint8_t)(asm_ || lsm);
  int ntop = up ? (maxX + 1 - x < size ? maxX + 1 - x : size) : 0;
  int ntr = (up && have_ar) ? maxX + 1 - (x + size) : 0;
  ntr = ntr < 0 ? 0 : (ntr > size ? size : ntr);
  int nl = left ? (maxY + 1 - y < size ? maxY + 1 - y : size) : 0;
  int nbl = (left && have_bl) ? maxY + 1 - (y + size) : 0;
  nbl = nbl < 0 ? 0 : (nbl > size ? size : nbl);
  jb.n_top_px = (int16_t)ntop, jb.n_topright_px = (int16_t)ntr, jb.n_left_px = (int16_t)nl,
  jb.n_bottomleft_px = (int16_t)nbl;
  int32_t *pred = (int32_t *)malloc(sizeof(int32_t) * size * size);
  or_intra_predict(d->frame, d->W, d->bd, &jb, pred);
  int32_t q[1024];
  int eob = 0, tx_type = 0, dcat = 0;
  const int np = size < 32 ? size * size : 1024;
  memset(q, 0, sizeof(q));
  if (!skip) {
    const int dir = fi < 5 ? FI_DIR[fi] : mode;
    eob = read_coeffs(d, t, c, r, size, dir, &tx_type, q, &dcat);
    if (eob < 0) {
      free(pred);
      return -1;
    }
  }
  /* reconstruction */
  const int txs = tx_size_of(size);
  int32_t *res = (int32_t *)calloc((size_t)size * size, sizeof(int32_t));
  if (eob > 0) {
    const int dq_shift = size == 64 ? 2 : (size == 32 ? 1 : 0);
    const int32_t dcq = d->bd > 8 ? DC_QLOOKUP_10[d->qidx] : DC_QLOOKUP_8[d->qidx];
    const int32_t acq = d->bd > 8 ? AC_QLOOKUP_10[d->qidx] : AC_QLOOKUP_8[d->qidx];
    const int n = size < 32 ? size : 32;
    const int64_t lo = -(1ll << (7 + d->bd)), hi = (1ll << (7 + d->bd)) - 1;
    int32_t *dq = (int32_t *)calloc((size_t)size * size, sizeof(int32_t));
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const int32_t v = q[i * n + j];
        if (!v) continue;
        const int64_t a = v < 0 ? -(int64_t)v : v;
        int64_t x2 = (a * (i == 0 && j == 0 ? dcq : acq)) & 0xFFFFFF;
        x2 >>= dq_shift;
        if (v < 0) x2 = -x2;
        dq[i * size + j] = (int32_t)(x2 < lo ? lo : (x2 > hi ? hi : x2));
      }
    or_inv_txfm2d(dq, res, size, txs, tx_type, d->bd);
    free(dq);
  }
  const int mx = (1 << d->bd) - 1;
  for (int i = 0; i < size; i++)
    for (int j = 0; j < size; j++) {
      int v = pred[i * size + j] + res[i * size + j];
      st(d, x + j, y + i, v < 0 ? 0 : (v > mx ? mx : v));
    }
  free(pred);
  free(res);
  /* contexts and the decoded flags */
  for (int k = 0; k < n4; k++) {
    t->above_dc[c + k] = (uint8_t)dcat;
    t->left_dc[r + k] = (uint8_t)dcat;
  }
  for (int i = 0; i < n4; i++)
    for (int j = 0; j < n4; j++) {
      d->ymode[(r + i) * W + c + j] = (uint8_t)mode;
      d->skip[(r + i) * W + c + j] = (uint8_t)skip;
      d->wlog[(r + i) * W + c + j] = (uint8_t)floor_log2((uint32_t)n4);
      t->bd_flags[sr + i + 1][scl + j + 1] = 1;
    }
  or_dec_block b;
  memset(&b, 0, sizeof(b));
  b.x = x, b.y = y, b.size = (uint8_t)size, b.mode = (uint8_t)mode, b.delta = (int8_t)delta, b.fi_mode = (uint8_t)fi;
  b.skip = (uint8_t)skip, b.tx_type = (uint8_t)tx_type, b.eob = eob;
  return push_block(d, &b, q, skip ? 0 : np);
}

static int decode_partition(Dec *d, Tile *t, int r, int c, int bs4) {
  if (r >= d->mi_rows || c >= d->mi_cols) return 0;
  const int W = d->mi_cols;
  const int up = r > t->mr0, left = c > t->mc0;
  const int bsl = floor_log2((uint32_t)bs4);
  const int half = bs4 / 2;
  const int has_rows = r + half < d->mi_rows, has_cols = c + half < d->mi_cols;
  int part;
  if (bs4 < 2) {
    part = 0;
  } else if (has_rows && has_cols) {
    const int above = up && d->wlog[(r - 1) * W + c] < bsl, lft = left && d->wlog[r * W + c - 1] < bsl;
    const int ctx = lft * 2 + above;
    part = read_symbol(&t->sd, t->cdf.part[(bsl - 1) * 4 + ctx], bsl == 1 ? 4 : (bsl == 5 ? 8 : 10), 1);
  } else {
    return -2; /* partial superblocks are outside the subset */
  }
  if (part == 0) return decode_block(d, t, r, c, bs4 * 4);
  if (part != 3) return -2;
  int rc;
  if ((rc = decode_partition(d, t, r, c, half))) return rc;
  if ((rc = decode_partition(d, t, r, c + half, half))) return rc;
  if ((rc = decode_partition(d, t, r + half, c, half))) return rc;
  return decode_partition(d, t, r + half, c + half, half);
}

static int tile_log2(int blk, int target) {
  int k = 0;
  while ((blk << k) < target) k++;
  return k;
}

int or_av1_decode_still(const uint8_t *data, size_t size, void *frame, int max_w, int max_h, or_dec_info *info,
                        or_dec_block *blocks, int max_blocks, int32_t *levels, size_t max_levels) {
  Dec d;
  memset(&d, 0, sizeof(d));
  size_t i = 0;
  int have_seq = 0, done = 0;
  int frame_w_bits = 0, frame_h_bits = 0, W = 0, H = 0;
  while (i < size && !done) {
    const uint8_t h = data[i++];
    const int type = (h >> 3) & 15, ext = (h >> 2) & 1, has_size = (h >> 1) & 1;
    if (ext) i++;
    if (!has_size) return -2;
    const size_t sz = leb128(data, size, &i);
    if (i + sz > size) return -1;
    BR br = {data + i, sz, 0};
    if (type == 1) { /* sequence header */
      if (f(&br, 3) != 0 || f(&br, 1) != 1 || f(&br, 1) != 1) return -2; /* profile 0, reduced still picture */
      (void)f(&br, 5);
      frame_w_bits = (int)f(&br, 4) + 1, frame_h_bits = (int)f(&br, 4) + 1;
      W = (int)f(&br, frame_w_bits) + 1, H = (int)f(&br, frame_h_bits) + 1;
      d.sb128 = (int)f(&br, 1);
      if (f(&br, 1) != 1 || f(&br, 1) != 1) return -2; /* filter intra, intra edge filter on */
      if (f(&br, 1) || f(&br, 1) || f(&br, 1)) return -2; /* superres, cdef, restoration off */
      d.bd = f(&br, 1) ? 10 : 8;
      if (f(&br, 1) != 1) return -2; /* monochrome */
      if (f(&br, 1)) return -2;      /* color description */
      (void)f(&br, 1);               /* color_range */
      if (f(&br, 1)) return -2;      /* film grain */
      have_seq = 1;
    } else if (type == 6 && have_seq) { /* frame OBU */
      if (W > max_w || H > max_h || W % 4 || H % 4) return -1;
      d.W = W, d.H = H, d.mi_cols = W / 4, d.mi_rows = H / 4;
      if (f(&br, 1)) return -2; /* disable_cdf_update */
      if (f(&br, 1)) return -2; /* allow_screen_content_tools */
      if (f(&br, 1)) return -2; /* render size */
      /* tile_info */
      const int sb_shift = d.sb128 ? 5 : 4, sb_size = sb_shift + 2;
      const int sb_cols = (d.mi_cols + (1 << sb_shift) - 1) >> sb_shift;
      const int sb_rows = (d.mi_rows + (1 << sb_shift) - 1) >> sb_shift;
      const int max_tw = 4096 >> sb_size, max_ta = (4096 * 2304) >> (2 * sb_size);
      const int min_c = tile_log2(max_tw, sb_cols);
      const int max_c = tile_log2(1, sb_cols < 64 ? sb_cols : 64), max_r = tile_log2(1, sb_rows < 64 ? sb_rows : 64);
      const int min_t = min_c > tile_log2(max_ta, sb_rows * sb_cols) ? min_c : tile_log2(max_ta, sb_rows * sb_cols);
      if (!f(&br, 1)) return -2; /* uniform spacing only */
      int cl = min_c;
      while (cl < max_c) {
        if (f(&br, 1)) cl++;
        else break;
      }
      const int tw = (sb_cols + (1 << cl) - 1) >> cl;
      int k = 0;
      for (int s = 0; s < sb_cols; s += tw) d.col_starts[k++] = s << sb_shift;
      d.col_starts[k] = d.mi_cols;
      d.tile_cols = k;
      int rl = min_t - cl > 0 ? min_t - cl : 0;
      while (rl < max_r) {
        if (f(&br, 1)) rl++;
        else break;
      }
      const int th = (sb_rows + (1 << rl) - 1) >> rl;
      k = 0;
      for (int s = 0; s < sb_rows; s += th) d.row_starts[k++] = s << sb_shift;
      d.row_starts[k] = d.mi_rows;
      d.tile_rows = k;
      int tsb = 4;
      if (cl > 0 || rl > 0) {
        (void)f(&br, rl + cl);
        tsb = (int)f(&br, 2) + 1;
      }
      d.qidx = (int)f(&br, 8);
      if (f(&br, 1)) return -2; /* DeltaQYDc */
      if (f(&br, 1)) return -2; /* using_qmatrix */
      if (f(&br, 1)) return -2; /* segmentation */
      if (d.qidx > 0 && f(&br, 1)) return -2; /* delta q */
      if (d.qidx == 0) return -2;              /* lossless is outside the subset */
      if (f(&br, 6) || f(&br, 6)) return -2;   /* loop filter levels 0 */
      (void)f(&br, 3);
      if (f(&br, 1)) return -2; /* loop filter deltas */
      if (f(&br, 1)) return -2; /* tx_mode_select: TX_MODE_LARGEST only */
      if (f(&br, 1)) return -2; /* reduced_tx_set */
      byte_align(&br);
      const int n_tiles = d.tile_cols * d.tile_rows;
      if (n_tiles > 1) {
        if (f(&br, 1)) return -2; /* tile_start_and_end_present_flag */
        byte_align(&br);
      }
      size_t p = br.pos / 8;
      const size_t mi = (size_t)d.mi_cols * d.mi_rows;
      d.ymode = calloc(mi, 1), d.skip = calloc(mi, 1), d.wlog = calloc(mi, 1);
      d.frame = frame;
      d.blocks = blocks, d.max_blocks = max_blocks, d.levels = levels, d.max_levels = max_levels;
      memset(frame, 0, (size_t)W * H * (d.bd > 8 ? 2 : 1));
      Tile *t = (Tile *)malloc(sizeof(Tile));
      int rc = 0;
      for (int tn = 0; tn < n_tiles && !rc; tn++) {
        size_t tsz;
        if (tn + 1 < n_tiles) {
          tsz = 0;
          for (int b = 0; b < tsb; b++) tsz |= (size_t)data[i + p + b] << (8 * b);
          tsz += 1;
          p += (size_t)tsb;
        } else {
          tsz = sz - p;
        }
        if (p + tsz > sz) {
          rc = -1;
          break;
        }
        memset(t, 0, sizeof(Tile));
        t->mc0 = d.col_starts[tn % d.tile_cols], t->mc1 = d.col_starts[tn % d.tile_cols + 1];
        t->mr0 = d.row_starts[tn / d.tile_cols], t->mr1 = d.row_starts[tn / d.tile_cols + 1];
        cdf_init(&t->cdf, d.qidx);
        init_symbol(&t->sd, data + i + p, tsz);
        const int sb4 = d.sb128 ? 32 : 16;
        for (int r = t->mr0; r < t->mr1 && !rc; r += sb4) {
          memset(t->left_dc, 0, sizeof(t->left_dc));
          for (int c = t->mc0; c < t->mc1 && !rc; c += sb4) {
            /* clear_block_decoded_flags */
            const int sbw4 = t->mc1 - c, sbh4 = t->mr1 - r;
            for (int y = -1; y <= sb4; y++)
              for (int x = -1; x <= sb4; x++)
                t->bd_flags[y + 1][x + 1] = (uint8_t)((y < 0 && x < sbw4) || (x < 0 && y < sbh4));
            t->bd_flags[sb4 + 1][0] = 0;
            t->sb_r = r, t->sb_c = c;
            rc = decode_partition(&d, t, r, c, sb4);
          }
        }
        p += tsz;
      }
      free(t);
      free(d.ymode), free(d.skip), free(d.wlog);
      if (rc) return rc;
      done = 1;
    }
    i += sz;
  }
  if (!done) return -1;
  if (info) {
    info->width = W, info->height = H, info->bd = d.bd, info->sb = d.sb128 ? 128 : 64, info->qindex = d.qidx;
    info->tile_cols = d.tile_cols, info->tile_rows = d.tile_rows, info->n_blocks = d.n_blocks;
    info->n_levels = (int64_t)d.n_levels;
  }
  return 0;
}
