/* Oracle (test infrastructure only): coefficient rate tables and the intra
 * tx-type rate from the AV1 default CDFs.
 *
 *   or_cost_symbol       context.rs:388 av1_cost_symbol
 *   syntax_rate          enc/quant.rs:494 syntax_rate_from_cdf
 *   or_build_coeff_costs enc/quant.rs:514 build_coeff_cost_tables
 *                        (-> :523 build_coeff_cost_tables_from_fc over
 *                        coeff_c.rs:651 CoeffFc::default_for_qindex)
 *   or_intra_txt_rate    leaf_funnel/rate_tables.rs:291 MdRates::txt_rate
 *                        (intra arm), ext-tx sets coeff_c.rs:153-212
 * Paths are relative to rust/crates/svtav1-{entropy,encoder}/src.  The CDF
 * data is av1_coef_cdfs.h (generated from the reference's default_cdfs.rs
 * by tests/golden/make_coef_cdfs.py). */
#include <string.h>

#include "av1_coef_cdfs.h"
#include "zav1_oracle.h"

/* default_cdfs.rs:19 coef_q_ctx */
int or_coef_q_ctx(int base_qindex) {
  if (base_qindex <= 20) return 0;
  if (base_qindex <= 60) return 1;
  if (base_qindex <= 120) return 2;
  return 3;
}

int or_cost_symbol(unsigned p15) {
  if (p15 < 1) p15 = 1;
  if (p15 > 32767) p15 = 32767;
  int msb = 31 - __builtin_clz(p15);
  int shift = 14 - msb;
  unsigned prob = (((p15 << shift) * 256u) + (1u << 14)) >> 15;
  if (prob < 1) prob = 1;
  if (prob > 255) prob = 255;
  return AV1_PROB_COST[prob - 128] + shift * 512;
}

/* Costs of the symbols of one inverse CDF; stops after the symbol whose
 * stored value is 0 (the last one).  Probabilities floored at 4. */
void or_syntax_rate(int32_t *cost, const uint16_t *icdf, int max_syms) {
  unsigned below = 0;
  for (int i = 0; i < max_syms; i++) {
    unsigned cum = 32768u - icdf[i];
    unsigned p = cum - below;
    if (p < 4) p = 4;
    below = cum;
    cost[i] = or_cost_symbol(p);
    if (icdf[i] == 0) break;
  }
}

void or_build_coeff_costs(int base_qindex, or_txb_costs txb[10], or_eob_costs eob[7][2]) {
  const int q = or_coef_q_ctx(base_qindex);
  memset(txb, 0, sizeof(or_txb_costs) * 10);
  memset(eob, 0, sizeof(or_eob_costs) * 14);
  for (int plane = 0; plane < 2; plane++)
    for (int ctx = 0; ctx < 2; ctx++) {
      or_syntax_rate(eob[0][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF16[q][plane][ctx], 6);
      or_syntax_rate(eob[1][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF32[q][plane][ctx], 7);
      or_syntax_rate(eob[2][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF64[q][plane][ctx], 8);
      or_syntax_rate(eob[3][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF128[q][plane][ctx], 9);
      or_syntax_rate(eob[4][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF256[q][plane][ctx], 10);
      or_syntax_rate(eob[5][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF512[q][plane][ctx], 11);
      or_syntax_rate(eob[6][plane].eob_cost[ctx], AV1_EOB_FLAG_CDF1024[q][plane][ctx], 12);
    }
  for (int t = 0; t < 5; t++)
    for (int plane = 0; plane < 2; plane++) {
      or_txb_costs *c = &txb[t * 2 + plane];
      for (int k = 0; k < 13; k++) or_syntax_rate(c->txb_skip_cost[k], AV1_TXB_SKIP_CDF[q][t][k], 3);
      for (int k = 0; k < 4; k++) or_syntax_rate(c->base_eob_cost[k], AV1_COEFF_BASE_EOB_CDF[q][t][plane][k], 4);
      for (int k = 0; k < 42; k++) {
        int32_t *b = c->base_cost[k];
        or_syntax_rate(b, AV1_COEFF_BASE_CDF[q][t][plane][k], 5);
        /* derived entries: [4] unused, [5] level-1 cost incl. the sign bit
         * relative to level 0, [6] / [7] steps 1->2 and 2->3 */
        b[4] = 0;
        b[5] = b[1] + 512 - b[0];
        b[6] = b[2] - b[1];
        b[7] = b[3] - b[2];
      }
      for (int k = 0; k < 9; k++) or_syntax_rate(c->eob_extra_cost[k], AV1_EOB_EXTRA_CDF[q][t][plane][k], 3);
      for (int k = 0; k < 3; k++) or_syntax_rate(c->dc_sign_cost[k], AV1_DC_SIGN_CDF[q][plane][k], 3);
      const int tb = t < 3 ? t : 3;
      for (int k = 0; k < 21; k++) {
        int32_t br[5] = {0};
        or_syntax_rate(br, AV1_COEFF_BR_CDF[q][tb][plane][k], 5);
        int32_t *l = c->lps_cost[k];
        /* lps_cost[i] for base-range increments i = 0..12: every group of
         * three continues with one more br symbol; the fourth br symbol
         * (continue) carries into the next group */
        int32_t carry = 0;
        for (int i = 0; i < 12; i++) {
          l[i] = carry + br[i % 3];
          if (i % 3 == 2) carry += br[3];
        }
        l[12] = carry;
        l[13] = l[0];
        for (int i = 1; i <= 12; i++) l[13 + i] = l[i] - l[i - 1];
      }
    }
}

/* ext-tx set machinery, intra only (coeff_c.rs:153-212) */
static const int SQR[19] = {0, 1, 2, 3, 4, 0, 0, 1, 1, 2, 2, 3, 3, 0, 0, 1, 1, 2, 2};
static const int SQR_UP[19] = {0, 1, 2, 3, 4, 1, 1, 2, 2, 3, 3, 4, 4, 2, 2, 3, 3, 4, 4};
static const int NUM_EXT_TX_SET[6] = {1, 2, 5, 7, 12, 16};
static const int EXT_TX_SET_INDEX_INTRA[6] = {0, -1, 2, 1, -1, -1};
static const unsigned char EXT_TX_USED[6][16] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, {1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0},
    {1, 1, 1, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 0},
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}};
static const unsigned char EXT_TX_IND[6][16] = {
    {0}, {1, 0}, {1, 3, 4, 2}, {1, 5, 6, 4, 0, 0, 0, 0, 0, 0, 2, 3}, {3, 4, 5, 8, 6, 7, 9, 10, 11, 0, 1, 2},
    {7, 8, 9, 12, 10, 11, 13, 14, 15, 0, 1, 2, 3, 4, 5, 6}};

/* get_ext_tx_set_type(tx_size, is_inter = 0, reduced = 0) */
int or_intra_ext_tx_set_type(int tx_size) {
  if (SQR_UP[tx_size] > 3) return 0;
  if (SQR_UP[tx_size] == 3) return 0;
  return SQR[tx_size] == 2 ? 2 : 3;
}

int or_intra_tx_type_used(int tx_size, int tx_type) {
  return EXT_TX_USED[or_intra_ext_tx_set_type(tx_size)][tx_type];
}

int or_intra_txt_rate(int tx_size, int intra_dir, int tx_type) {
  const int set = or_intra_ext_tx_set_type(tx_size);
  if (NUM_EXT_TX_SET[set] <= 1) return 0;
  const int eset = EXT_TX_SET_INDEX_INTRA[set];
  if (eset <= 0) return 0;
  if (!EXT_TX_USED[set][tx_type]) return 0;
  int32_t cost[17] = {0};
  or_syntax_rate(cost, AV1_INTRA_EXT_TX_CDF[eset][SQR[tx_size]][intra_dir], 17);
  return cost[EXT_TX_IND[set][tx_type]];
}

/* BLOCK_SIZES index of a w x h block (4x4 .. 128x128, then the 1:4 sizes) */
int or_bsize_index(int w, int h) {
  static const int W[22] = {4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 4, 16, 8, 32, 16, 64};
  static const int H[22] = {4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 64, 32, 64, 128, 64, 128, 16, 4, 32, 8, 64, 16};
  for (int i = 0; i < 22; i++)
    if (W[i] == w && H[i] == h) return i;
  return -1;
}

/* Luma rate (1/512 bit) of an intra candidate of a key frame from the
 * default CDFs (leaf_funnel/rate_tables.rs:146 build_md_rates, as the fast
 * loop adds it, mod.rs:2060-2185): the kf y-mode symbol under the above /
 * left mode contexts, the angle delta of a directional mode on blocks of at
 * least 8x8, and for DC_PRED on blocks up to 32x32 the filter-intra flag
 * (+ its mode). */
int or_intra_mode_rate(int w, int h, int actx, int lctx, int mode, int delta, int fi) {
  int32_t c[16] = {0};
  or_syntax_rate(c, AV1_KF_Y_CDF[actx][lctx], 13);
  int r = c[mode];
  if (mode >= 1 && mode <= 8 && w >= 8 && h >= 8) {
    int32_t a[8] = {0};
    or_syntax_rate(a, AV1_ANGLE_DELTA_CDF[mode - 1], 7);
    r += a[delta + 3];
  }
  if (mode == 0 && w <= 32 && h <= 32) {
    int32_t f[2] = {0}, m[5] = {0};
    or_syntax_rate(f, AV1_FILTER_INTRA_CDF[or_bsize_index(w, h)], 2);
    r += f[fi != 5];
    if (fi != 5) {
      or_syntax_rate(m, AV1_FILTER_INTRA_MODE_CDF, 5);
      r += m[fi];
    }
  }
  return r;
}
