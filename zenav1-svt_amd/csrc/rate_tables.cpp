// Host side: coefficient rate tables and the intra tx-type rate from the AV1
// default CDFs.  No device work.
//
//   zav1_build_coeff_costs  rust/crates/svtav1-encoder/src/quant.rs:514
//                           build_coeff_cost_tables (-> :523 ..._from_fc over
//                           svtav1-entropy/src/coeff_c.rs:651 default_for_qindex)
//   zav1_intra_txt_rate     rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:291
//                           MdRates::txt_rate, intra arm (ext-tx sets
//                           svtav1-entropy/src/coeff_c.rs:153-212)
// Symbol costs follow av1_cost_symbol (svtav1-entropy/src/context.rs:388).
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>

#include "../../include/zav1_dsp.h"
#include "av1_coef_cdfs.h"
#include "av1_part_cdfs.h"

namespace {

// Cost in 1/512 bit of a symbol of probability p15 / 32768: normalise p15 into
// [2^14, 2^15), round it to an 8-bit probability, table lookup + whole bits.
int symbol_cost(uint32_t p15) {
  p15 = p15 < 1 ? 1 : (p15 > 32767 ? 32767 : p15);
  const int whole = __builtin_clz(p15) - 17;  // 14 - msb
  uint32_t p8 = ((p15 << whole) * 256u + 16384u) >> 15;
  p8 = p8 < 1 ? 1 : (p8 > 255 ? 255 : p8);
  return AV1_PROB_COST[p8 - 128] + 512 * whole;
}

// Per-symbol costs of one stored inverse CDF: symbol k's probability is the
// step between consecutive stored values (floored at 4); the symbol whose
// stored value is 0 is the last.
template <size_t K>
void cdf_costs(int32_t* out, const uint16_t (&icdf)[K], int n_max) {
  uint32_t prev = 32768;
  for (int k = 0; k < n_max; k++) {
    const uint32_t step = prev - icdf[k];
    out[k] = symbol_cost(step < 4 ? 4 : step);
    prev = icdf[k];
    if (icdf[k] == 0) return;
  }
}

int q_context(int qindex) { return qindex > 120 ? 3 : qindex > 60 ? 2 : qindex > 20 ? 1 : 0; }

// intra ext-tx set type per tx size (DCT only from 32 points; 16x16 the
// 5-type set, every other size the 7-type set), the set's tx types and each
// type's symbol index within its set
constexpr uint8_t kSqr[19] = {0, 1, 2, 3, 4, 0, 0, 1, 1, 2, 2, 3, 3, 0, 0, 1, 1, 2, 2};
constexpr uint8_t kSqrUp[19] = {0, 1, 2, 3, 4, 1, 1, 2, 2, 3, 3, 4, 4, 2, 2, 3, 3, 4, 4};
int intra_set(int tx_size) { return kSqrUp[tx_size] >= 3 ? 0 : (kSqr[tx_size] == 2 ? 2 : 3); }
// set 2 (DTT4_IDTX): DCT_DCT, ADST_DCT, DCT_ADST, ADST_ADST, IDTX;
// set 3 (DTT4_IDTX_1DDCT): those and V_DCT, H_DCT
constexpr int8_t kSym2[16] = {1, 3, 4, 2, -1, -1, -1, -1, -1, 0, -1, -1, -1, -1, -1, -1};
constexpr int8_t kSym3[16] = {1, 5, 6, 4, -1, -1, -1, -1, -1, 0, 2, 3, -1, -1, -1, -1};

}  // namespace

extern "C" {

int zav1_build_coeff_costs(int base_qindex, zav1_coeff_costs* out) {
  if (base_qindex < 0 || base_qindex > 255 || !out) return ZAV1_EINVAL;
  const int q = q_context(base_qindex);
  std::memset(out, 0, sizeof(*out));
  for (int p = 0; p < 2; p++)
    for (int c = 0; c < 2; c++) {
      cdf_costs(out->eob[0][p].eob_cost[c], AV1_EOB_FLAG_CDF16[q][p][c], 6);
      cdf_costs(out->eob[1][p].eob_cost[c], AV1_EOB_FLAG_CDF32[q][p][c], 7);
      cdf_costs(out->eob[2][p].eob_cost[c], AV1_EOB_FLAG_CDF64[q][p][c], 8);
      cdf_costs(out->eob[3][p].eob_cost[c], AV1_EOB_FLAG_CDF128[q][p][c], 9);
      cdf_costs(out->eob[4][p].eob_cost[c], AV1_EOB_FLAG_CDF256[q][p][c], 10);
      cdf_costs(out->eob[5][p].eob_cost[c], AV1_EOB_FLAG_CDF512[q][p][c], 11);
      cdf_costs(out->eob[6][p].eob_cost[c], AV1_EOB_FLAG_CDF1024[q][p][c], 11);
    }
  for (int ts = 0; ts < 5; ts++)
    for (int p = 0; p < 2; p++) {
      zav1_txb_costs& t = out->txb[ts * 2 + p];
      for (int c = 0; c < 13; c++) cdf_costs(t.txb_skip_cost[c], AV1_TXB_SKIP_CDF[q][ts][c], 2);
      for (int c = 0; c < 4; c++) cdf_costs(t.base_eob_cost[c], AV1_COEFF_BASE_EOB_CDF[q][ts][p][c], 3);
      for (int c = 0; c < 42; c++) {
        int32_t* b = t.base_cost[c];
        cdf_costs(b, AV1_COEFF_BASE_CDF[q][ts][p][c], 4);
        b[4] = 0;
        b[5] = b[1] - b[0] + 512;  // level 1 vs 0, with the sign bit
        b[6] = b[2] - b[1];
        b[7] = b[3] - b[2];
      }
      for (int c = 0; c < 9; c++) cdf_costs(t.eob_extra_cost[c], AV1_EOB_EXTRA_CDF[q][ts][p][c], 2);
      for (int c = 0; c < 3; c++) cdf_costs(t.dc_sign_cost[c], AV1_DC_SIGN_CDF[q][p][c], 2);
      for (int c = 0; c < 21; c++) {
        std::array<int32_t, 4> br{};
        cdf_costs(br.data(), AV1_COEFF_BR_CDF[q][ts < 3 ? ts : 3][p][c], 4);
        int32_t* l = t.lps_cost[c];
        // the base range is coded in up to four br symbols of three
        // increments each; br[3] continues to the next symbol
        for (int k = 0; k <= 12; k++) l[k] = (k / 3) * br[3] + (k < 12 ? br[k % 3] : 0);
        l[13] = l[0];
        for (int k = 1; k <= 12; k++) l[13 + k] = l[k] - l[k - 1];
      }
    }
  return ZAV1_OK;
}

int zav1_intra_txt_rate(int tx_size, int intra_dir, int tx_type) {
  if (tx_size < 0 || tx_size > 18 || intra_dir < 0 || intra_dir > 12 || tx_type < 0 || tx_type > 15)
    return ZAV1_EINVAL;
  const int set = intra_set(tx_size);
  if (set == 0) return 0;
  const int sym = set == 2 ? kSym2[tx_type] : kSym3[tx_type];
  if (sym < 0) return 0;
  // ext_tx_set_index for intra: set 2 -> 2, set 3 -> 1
  const int eset = set == 2 ? 2 : 1;
  int32_t cost[16] = {0};
  cdf_costs(cost, AV1_INTRA_EXT_TX_CDF[eset][kSqr[tx_size]][intra_dir], 16);
  return cost[sym];
}

}  // extern "C"

extern "C" int zav1_intra_mode_rate(int w, int h, int above_ctx, int left_ctx, int mode, int delta, int fi_mode) {
  // rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:146 build_md_rates,
  // key-frame luma terms as the fast loop sums them (mod.rs:2060-2185)
  static const int kBw[22] = {4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 4, 16, 8, 32, 16, 64};
  static const int kBh[22] = {4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 64, 32, 64, 128, 64, 128, 16, 4, 32, 8, 64, 16};
  int bsize = -1;
  for (int i = 0; i < 22 && bsize < 0; i++)
    if (kBw[i] == w && kBh[i] == h) bsize = i;
  if (bsize < 0 || above_ctx < 0 || above_ctx > 4 || left_ctx < 0 || left_ctx > 4 || mode < 0 || mode > 12 ||
      delta < -3 || delta > 3 || fi_mode < 0 || fi_mode > 5 || (fi_mode < 5 && (mode != 0 || w > 32 || h > 32)))
    return ZAV1_EINVAL;
  int32_t y[16] = {0};
  cdf_costs(y, AV1_KF_Y_CDF[above_ctx][left_ctx], 13);
  int rate = y[mode];
  const bool directional = mode >= 1 && mode <= 8;
  if (directional && w >= 8 && h >= 8) {
    int32_t a[8] = {0};
    cdf_costs(a, AV1_ANGLE_DELTA_CDF[mode - 1], 7);
    rate += a[delta + 3];
  }
  if (mode == 0 && w <= 32 && h <= 32) {  // the use_filter_intra flag (+ its mode)
    int32_t flag[2] = {0};
    cdf_costs(flag, AV1_FILTER_INTRA_CDF[bsize], 2);
    rate += flag[fi_mode == 5 ? 0 : 1];
    if (fi_mode != 5) {
      int32_t m[5] = {0};
      cdf_costs(m, AV1_FILTER_INTRA_MODE_CDF, 5);
      rate += m[fi_mode];
    }
  }
  return rate;
}

// The key-frame full mode-decision lambda of an allintra still
// (rust/crates/svtav1-encoder/src/pd0.rs:223 kf_full_lambda_8bit and :294
// kf_full_lambda_bd10): rdmult (3.3 + 0.0015 q) q^2 from the frame's DC
// quantizer, then at 8 bits the key-frame factor 150/128, at 10 bits the
// rounding >> 4 and factor 128/128; the PSNR lambda weight (150 from CLI qp
// 16, 175 from 56; none below 16) / 128; x16 at 10 bits.
#include "av1_qlookup_tables.h"
extern "C" uint64_t zav1_kf_full_lambda(int qindex, int cli_qp, int bd) {
  if (qindex < 0 || qindex > 255 || (bd != 8 && bd != 10)) return 0;
  const double q = bd == 10 ? DC_QLOOKUP_10[qindex] : DC_QLOOKUP_8[qindex];
  int64_t rdmult = (int64_t)((3.3 + 0.0015 * q) * q * q);
  if (bd == 10) rdmult = std::max<int64_t>((rdmult + 8) >> 4, 1);
  rdmult = (rdmult * (bd == 10 ? 128 : 150)) >> 7;
  uint64_t lambda = (uint32_t)rdmult;
  const uint64_t weight = cli_qp >= 56 ? 175 : (cli_qp >= 16 ? 150 : 0);
  if (weight) lambda = (lambda * weight) >> 7;
  return bd == 10 ? lambda * 16 : lambda;
}

extern "C" int zav1_pd0_rates(zav1_pd0_cfg* c) {
  if (!c || c->qindex < 0 || c->qindex > 255 || c->qp < 0 || c->qp > 63 || c->width <= 0 || c->height <= 0)
    return ZAV1_EINVAL;
  // PARTITION_SPLIT (symbol 3) of each square size's first context row; the
  // 8x8 row has 4 symbols, PARTITION_NONE read from it for every size
  for (int s = 0; s < 4; s++) {
    int32_t cost[10] = {};
    const uint16_t(&row)[11] = AV1_PARTITION_CDF[s * 4];
    cdf_costs(cost, row, s == 0 ? 4 : 10);
    c->split_bits[s] = cost[3];
    if (s == 0) c->none_bits = cost[0];
  }
  int32_t sk[2] = {};
  cdf_costs(sk, AV1_SKIP_CDF[0], 2);
  c->skip0_bits = sk[0];
  for (int t = 0; t < 3; t++) c->txt_dc[t] = zav1_intra_txt_rate(t, 0, 0);
  c->lambda = zav1_kf_full_lambda(c->qindex, c->qp, 8);
  // input_resolution_factor (pd0.rs:371) of the picture's pixel count
  const long px = (long)c->width * c->height;
  static const int kFactor[7] = {0, 1, 2, 3, 4, 4, 4};
  const long th[6] = {0x28500, 0x4CE00, 0xA1400, 0x16DA00, 0x535200, 0x140A000};
  int r = 0;
  while (r < 6 && px >= th[r]) r++;
  c->ires = kFactor[r];
  return ZAV1_OK;
}
