// Coefficient-context and rate helpers shared by the RDOQ trellis, the
// coefficient rate and the partition pre-pass: the cost-table layout of one
// (txs_ctx, plane) TxbCosts (rust/crates/svtav1-encoder/src/quant.rs:459) and
// EobCosts (:469), the per-size constants, and the nz-map / base-range
// contexts of rust/crates/svtav1-entropy/src/coeff_c.rs over a level
// accessor lv(r, c) that returns 0 past the block's right / bottom edge.
#pragma once
#include <cstdint>

#include "zav1_dev.h"

namespace zav1 {
namespace cctx {
// one TxbCosts table (quant.rs:459) as 944 words, one EobCosts (:469) as 22
constexpr int kTxbWords = 944, kEobWords = 22;
constexpr int kSkip = 0, kBaseEob = 26, kBase = 38, kEobExtra = 374, kDcSign = 392, kLps = 398;
constexpr int kCls2D = 0, kClsHoriz = 1, kClsVert = 2;

constexpr int c_txw(int t) { return tx_w(t) > 32 ? 32 : tx_w(t); }
constexpr int c_txh(int t) { return tx_h(t) > 32 ? 32 : tx_h(t); }
constexpr int c_log2(int v) { return v <= 1 ? 0 : 1 + c_log2(v >> 1); }
constexpr int c_sqr(int t) { return c_log2(tx_w(t) < tx_h(t) ? tx_w(t) : tx_h(t)) - 2; }
constexpr int c_sqr_up(int t) { return c_log2(tx_w(t) > tx_h(t) ? tx_w(t) : tx_h(t)) - 2; }
// coeff_c.rs:119 txsize_entropy_ctx; :63 TXSIZE_LOG2_MINUS4, which follows
// the coded (32-capped) dimensions
constexpr int c_txs_ctx(int t) { return (c_sqr(t) + c_sqr_up(t) + 1) >> 1; }
constexpr int c_eob_ms(int t) { return c_log2(c_txw(t)) + c_log2(c_txh(t)) - 4; }
// quant.rs:238 TX_SCALE_TAB
constexpr int c_scale(int t) {
  return tx_w(t) * tx_h(t) > 1024 ? 2 : (tx_w(t) * tx_h(t) > 256 ? 1 : 0);
}

__device__ __forceinline__ int msb(uint32_t v) { return 31 - __clz(v); }
__device__ __forceinline__ int64_t rdcost(int64_t rdmult, int64_t rate, int64_t dist) {
  return ((rate * rdmult + 256) >> 9) + (dist << 7);
}
__device__ __forceinline__ int64_t coeff_dist(int32_t t, int32_t dq, int shift) {
  const int64_t d = ((int64_t)t - dq) * ((int64_t)1 << shift);
  return d * d;
}
__device__ __forceinline__ int golomb_cost(int a) {
  return a >= 15 ? 512 * (2 * (msb((uint32_t)(a - 14)) + 1) - 1) : 0;
}

// Coefficient contexts of a block of size TXS with class cls, over a level
// accessor lv(r, c) that returns 0 past the block's right / bottom edge.
// coeff_c.rs:406 nz_map_ctx (+ :306 nz_mag, :344 nz_map_ctx_offset_2d,
// :380 from_stats); si = scan index; the 2D offsets take their shape from
// the unadjusted size.
template <int TXS, class LV>
__device__ int nz_ctx(const LV& lv, int ci, int si, bool is_last, int cls) {
  constexpr int AW = c_txw(TXS), BWL = c_log2(AW), N = AW * c_txh(TXS);
  if (is_last) {
    if (si == 0) return 0;
    if (si <= N / 8) return 1;
    if (si <= N / 4) return 2;
    return 3;
  }
  const int r = ci >> BWL, c = ci & (AW - 1);
  auto c3 = [](int v) { return v < 3 ? v : 3; };
  int mag = c3(lv(r, c + 1)) + c3(lv(r + 1, c));
  if (cls == kCls2D) {
    mag += c3(lv(r + 1, c + 1)) + c3(lv(r, c + 2)) + c3(lv(r + 2, c));
  } else if (cls == kClsVert) {
    mag += c3(lv(r + 2, c)) + c3(lv(r + 3, c)) + c3(lv(r + 4, c));
  } else {
    mag += c3(lv(r, c + 2)) + c3(lv(r, c + 3)) + c3(lv(r, c + 4));
  }
  if ((cls | ci) == 0) return 0;
  const int ctx = min((mag + 1) >> 1, 4);
  if (cls == kCls2D) {
    if (ci == 0) return ctx;
    if (tx_w(TXS) < tx_h(TXS)) {
      if (r < 2) return ctx + 11;
    } else if (tx_w(TXS) > tx_h(TXS) && c < 2) {
      return ctx + 16;
    }
    return ctx + (r + c < 2 ? 1 : (r + c < 4 ? 6 : 21));
  }
  const int pos = cls == kClsHoriz ? c : r;
  return ctx + (pos == 0 ? 26 : (pos == 1 ? 31 : 36));
}
// coeff_c.rs:537 br_ctx
template <int TXS, class LV>
__device__ int br_ctx_of(const LV& lv, int ci, int cls) {
  constexpr int AW = c_txw(TXS), BWL = c_log2(AW);
  const int r = ci >> BWL, c = ci & (AW - 1);
  int mag = lv(r, c + 1) + lv(r + 1, c);
  mag += cls == kCls2D ? lv(r + 1, c + 1) : (cls == kClsHoriz ? lv(r, c + 2) : lv(r + 2, c));
  mag = min((mag + 1) >> 1, 6);
  if (ci == 0) return mag;
  if ((cls == kCls2D && r < 2 && c < 2) || (cls == kClsHoriz && c == 0) || (cls == kClsVert && r == 0))
    return mag + 7;
  return mag + 14;
}


// quant.rs:851 get_eob_cost (coeff_c.rs:606 eob_pos_token) over one TxbCosts
// (cost) and one EobCosts (ecost) table
__device__ inline int eob_cost_of(int eob, const int32_t* cost, const int32_t* ecost, int cls) {
  int t;
  if (eob < 33) {
    t = eob <= 2 ? eob : (eob <= 4 ? 3 : (eob <= 8 ? 4 : (eob <= 16 ? 5 : 6)));
  } else {
    const int e = min((eob - 1) >> 5, 16);
    t = e == 0 ? 6 : (e == 1 ? 7 : (e <= 3 ? 8 : (e <= 7 ? 9 : (e <= 15 ? 10 : 11))));
  }
  const int start = t == 0 ? 0 : (t == 1 ? 1 : (1 << (t - 2)) + 1);
  int c = ecost[(cls != kCls2D) * 11 + t - 1];
  if (t > 2) {
    const int cnt = t - 3, bit = ((eob - start) >> cnt) & 1;
    c += cost[kEobExtra + cnt * 2 + bit] + 512 * cnt;
  }
  return c;
}
}  // namespace cctx
}  // namespace zav1
