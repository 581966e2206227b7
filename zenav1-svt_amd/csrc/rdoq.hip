// RDOQ trellis for gfx950: svt_av1_optimize_b
// (rust/crates/svtav1-encoder/src/quant.rs:1312) on the path without
// quantization matrices, with the rate helpers of quant.rs:636-:906 and the
// coefficient contexts of rust/crates/svtav1-entropy/src/coeff_c.rs.
//
// The trellis is a sequential reverse scan per transform block: each step
// prices a coefficient against the levels of the coefficients after it,
// which earlier steps may have lowered.  Two forms, bit-identical:
//
// Lane form (large batches, caller workspace): one lane per block, 64
// blocks per workgroup.  A gather pass lays every block's (t, q) pairs out
// in scan order in the caller's workspace, with each block's packed level
// map, so a lane's walk reads consecutive pairs (one cache line serves
// eight steps) and many resident waves hide the loads.  A lane's level map
// lives in LDS as 4-bit levels min(|q|, 15), eight per dword, dwords
// interleaved across the 64 lanes.  Capping at 15 instead of the
// reference's 127 leaves every context unchanged: nz_mag clips each level at
// 3, and br_ctx's min((l1 + l2 + l3 + 1) >> 1, 6) saturates once the sum
// reaches 11.
//
// Wave form (small batches, where a few hundred blocks cannot keep the chip
// busy and the lane form's sequential walk is the whole latency): one wave
// per block, the block's coefficients and levels staged in LDS.  The wave
// runs the sequential part in lockstep, the eob phase (update_coeff_eob with
// its running rate and distortion, until five nonzero levels are kept) and
// the skip decision, stepping only at nonzero levels (the zeros between are
// priced across the lanes).  The rest of the walk, update_coeff_simple down to the
// cut-off, carries no running state: each position's level is a function
// of the final levels of the positions after it.  The wave solves that
// system as a fixpoint: every lane re-decides its positions against the
// current map until a pass changes nothing.  The dependencies run strictly
// from later to earlier scan positions, so the solution is unique and equals
// the sequential walk's, and every pass settles at least the latest
// unsettled position, so the passes end.  The wave then prices the DC
// position (update_coeff_general) when the walk reaches it.
//
// Group form (the latency form up to 256 coefficients): the wave form with
// 16 lanes per block and four blocks per wave, so one instruction stream
// advances four blocks' sequential eob phases (the phase that dominates:
// about twelve sequential steps per block on dense 16x16 candidates).
//
// Reads past the block's right or bottom edge return 0, which is what the
// reference's zero padding (4 columns, >= 6 rows) holds.  The cost tables of
// the launch's size (both plane types) and its three scans are staged in
// LDS once per workgroup.
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "coeff_ctx.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

namespace {
using namespace cctx;

// Level map and coefficient source of the lane form: 4-bit levels in
// interleaved LDS words, (t, q) pairs in scan order from the workspace
struct GatherMap {
  uint32_t* lv;      // the lane's word 0; word stride 64
  const int2* work;  // the block's pairs, scan order
  __device__ __forceinline__ int get(int p) const { return (lv[(p >> 3) * 64] >> ((p & 7) * 4)) & 15; }
  __device__ __forceinline__ void set(int p, int v) {
    uint32_t& w = lv[(p >> 3) * 64];
    const int sh = (p & 7) * 4;
    w = (w & ~(15u << sh)) | ((uint32_t)v << sh);
  }
  __device__ __forceinline__ int2 coef(int si, int) const { return work[si]; }
};
// ... of the wave form: byte levels and the raster block, all in LDS
// (QT: the staged quantizer levels' type -- 16 bits in the compact group
// form, where the caller guarantees they fit)
template <class QT>
struct WaveMapQ {
  uint8_t* lv;
  const int32_t* ts;
  const QT* qs;
  __device__ __forceinline__ int get(int p) const { return lv[p]; }
  __device__ __forceinline__ void set(int p, int v) { lv[p] = (uint8_t)v; }
  __device__ __forceinline__ int2 coef(int, int ci) const { return make_int2(ts[ci], (int32_t)qs[ci]); }
};
using WaveMap = WaveMapQ<int32_t>;

template <int TXS, class Map>
struct Trellis {
  static constexpr int AW = c_txw(TXS), AH = c_txh(TXS), BWL = c_log2(AW), N = AW * AH, NW = (N + 7) / 8;
  static constexpr int SHIFT = c_scale(TXS);
  Map m;
  int32_t* qc;         // levels and dequantized values, raster (written on change)
  int32_t* dq;
  const int32_t* cost; // TxbCosts of the block's plane type
  const int32_t* ecost;
  const uint16_t* scan;
  int64_t rdmult;
  int32_t dqv[2];
  int cls, dc_ctx;

  __device__ __forceinline__ int level(int r, int c) const {
    if (r >= AH || c >= AW) return 0;
    return m.get(r * AW + c);
  }
  __device__ __forceinline__ void set_level(int p, int v) { m.set(p, v < 15 ? v : 15); }
  __device__ __forceinline__ int lower_ctx(int ci, int si, bool is_last) const {
    return nz_ctx<TXS>([this](int r, int c) { return level(r, c); }, ci, si, is_last, cls);
  }
  // coeff_c.rs:456 br_ctx_eob
  __device__ __forceinline__ int br_ctx_eob(int ci) const {
    const int r = ci >> BWL, c = ci & (AW - 1);
    if (ci == 0) return 0;
    if ((cls == kCls2D && r < 2 && c < 2) || (cls == kClsHoriz && c == 0) || (cls == kClsVert && r == 0)) return 7;
    return 14;
  }
  __device__ __forceinline__ int br_ctx(int ci) const {
    return br_ctx_of<TXS>([this](int r, int c) { return level(r, c); }, ci, cls);
  }
  // nz_ctx (not last) and br_ctx of ci from one read of its five template
  // levels (the base-range template is the first three of the nz-map one in
  // every class), so the reads issue together
  __device__ __forceinline__ void ctx_pair(int ci, int* nzc, int* brc) const {
    const int r = ci >> BWL, c = ci & (AW - 1);
    const bool d2 = cls == kCls2D, hz = cls == kClsHoriz;
    const int r1 = d2 ? r + 1 : (hz ? r : r + 2), c1 = d2 ? c + 1 : (hz ? c + 2 : c);
    const int r2 = d2 ? r : (hz ? r : r + 3), c2 = d2 ? c + 2 : (hz ? c + 3 : c);
    const int r3 = d2 ? r + 2 : (hz ? r : r + 4), c3 = d2 ? c : (hz ? c + 4 : c);
    const int la = level(r, c + 1), lb = level(r + 1, c), l1 = level(r1, c1), l2 = level(r2, c2),
              l3 = level(r3, c3);
    const int mag = min(la, 3) + min(lb, 3) + min(l1, 3) + min(l2, 3) + min(l3, 3);
    const int bm = min((la + lb + l1 + 1) >> 1, 6);
    if (ci == 0) {
      *nzc = 0;  // (cls | ci) == 0 for the 2D class; the 1D classes at ci 0 below
      *brc = bm;
      if (!d2) *nzc = min((mag + 1) >> 1, 4) + 26;
      return;
    }
    const int ctx = min((mag + 1) >> 1, 4);
    if (d2) {
      if (tx_w(TXS) < tx_h(TXS) && r < 2) *nzc = ctx + 11;
      else if (tx_w(TXS) > tx_h(TXS) && c < 2) *nzc = ctx + 16;
      else *nzc = ctx + (r + c < 2 ? 1 : (r + c < 4 ? 6 : 21));
    } else {
      const int pos = hz ? c : r;
      *nzc = ctx + (pos == 0 ? 26 : (pos == 1 ? 31 : 36));
    }
    *brc = bm + (((d2 && r < 2 && c < 2) || (hz && c == 0) || (!d2 && !hz && r == 0)) ? 7 : 14);
  }
  __device__ __forceinline__ int base_cost(int ctx, int k) const { return cost[kBase + ctx * 8 + k]; }
  __device__ __forceinline__ int lps(int ctx, int k) const { return cost[kLps + ctx * 26 + k]; }
  __device__ __forceinline__ int br_cost(int level_, int ctx) const {
    return lps(ctx, min(level_ - 3, 12)) + golomb_cost(level_);
  }
  // quant.rs:758 get_coeff_cost_general
  __device__ int cost_general(bool is_last, int ci, int a, int sign, int ctx) const {
    int c = is_last ? cost[kBaseEob + ctx * 3 + min(a, 3) - 1] : base_cost(ctx, min(a, 3));
    if (a) {
      c += ci == 0 ? cost[kDcSign + dc_ctx * 2 + sign] : 512;
      if (a > 2) c += br_cost(a, is_last ? br_ctx_eob(ci) : br_ctx(ci));
    }
    return c;
  }
  // quant.rs:795 get_coeff_cost_eob
  __device__ int cost_eob(int ci, int a, int sign, int ctx) const {
    int c = cost[kBaseEob + ctx * 3 + min(a, 3) - 1];
    if (a) {
      c += ci == 0 ? cost[kDcSign + dc_ctx * 2 + sign] : 512;
      if (a > 2) c += br_cost(a, br_ctx_eob(ci));
    }
    return c;
  }
  // quant.rs:822 get_two_coeff_cost_simple (+ :736 br_cost_with_diff with
  // the two Golomb tables in closed form), in two halves: the level's cost
  // (b: its base-range context, for the second half), then the cost one
  // level lower, which only a coefficient that may be lowered needs
  __device__ int cost_simple(int ci, int a, int ctx, int* b) const {
    int c = base_cost(ctx, min(a, 3));
    if (a) {
      c += 512;
      if (a > 2) {
        *b = br_ctx(ci);
        c += lps(*b, min(a - 3, 12)) + golomb_cost(a);
      }
    }
    return c;
  }
  __device__ int cost_simple_low(int a, int ctx, int b, int c) const {
    int diff = a <= 3 ? base_cost(ctx, a + 4) : 0;
    if (a > 2) {
      if (a <= 15) diff += lps(b, min(a - 3, 12) + 13);
      if (a >= 15) {
        const int r = a - 14;
        diff += r == 1 ? 512 : ((r & (r - 1)) == 0 ? 1024 : 0);
      }
    }
    return c - diff;
  }
  // quant.rs:851 get_eob_cost
  __device__ int eob_cost(int eob) const { return eob_cost_of(eob, cost, ecost, cls); }
  // coefficient of scan index si; the dequantized value follows the level
  // as fp quantization sets it, sign((|q| * dequant) >> log_scale)
  __device__ __forceinline__ void fetch(int si, int& ci, int32_t& t, int32_t& q, int32_t& d) const {
    ci = scan[si];
    const int2 v = m.coef(si, ci);
    t = v.x, q = v.y;
    const int32_t a = (int32_t)(((int64_t)abs(q) * dqv[ci != 0]) >> SHIFT);
    d = q < 0 ? -a : a;
  }
  __device__ __forceinline__ void low(int a, int sign, int dv, int32_t* q, int32_t* d) const {
    const int lo = a - 1;
    *q = (-sign ^ lo) + sign;
    const int32_t dl = (int32_t)(((int64_t)lo * dv) >> SHIFT);
    *d = (-sign ^ dl) + sign;
  }

  // quant.rs:966 update_coeff_general
  __device__ void update_general(int* accu_rate, int64_t* accu_dist, int si, int eob) {
    int ci;
    int32_t t, qv, d;
    fetch(si, ci, t, qv, d);
    const int dv = dqv[ci != 0];
    const bool is_last = si == eob - 1;
    const int ctx = lower_ctx(ci, si, is_last);
    if (qv == 0) {
      *accu_rate += base_cost(ctx, 0);
      return;
    }
    const int sign = qv < 0, a = qv < 0 ? -qv : qv;
    const int64_t dist = coeff_dist(t, d, SHIFT), dist0 = coeff_dist(t, 0, SHIFT);
    const int rate = cost_general(is_last, ci, a, sign, ctx);
    const int64_t rd = rdcost(rdmult, rate, dist);
    int32_t q_low = 0, d_low = 0;
    int rate_low;
    int64_t dist_low;
    if (a == 1) {
      dist_low = dist0;
      rate_low = base_cost(ctx, 0);
    } else {
      low(a, sign, dv, &q_low, &d_low);
      dist_low = coeff_dist(t, d_low, SHIFT);
      rate_low = cost_general(is_last, ci, a - 1, sign, ctx);
    }
    if (rdcost(rdmult, rate_low, dist_low) < rd) {
      qc[ci] = q_low;
      dq[ci] = d_low;
      set_level(ci, a - 1);
      *accu_rate += rate_low;
      *accu_dist += dist_low - dist0;
    } else {
      *accu_rate += rate;
      *accu_dist += dist - dist0;
    }
  }
  // quant.rs:1056 update_coeff_simple's decision: the level kept at scan
  // index si (|q| or |q| - 1, q != 0), from the levels after it
  __device__ int simple_level(int ci, int si, int32_t t, int32_t qv) const {
    const int dv = dqv[ci != 0];
    const int ctx = lower_ctx(ci, si, false);
    const int a = qv < 0 ? -qv : qv;
    const int32_t at = abs(t), ad = (int32_t)(((int64_t)a * dv) >> SHIFT);
    int b = 0;
    const int rate = cost_simple(ci, a, ctx, &b);
    if (ad < at) return a;
    const int rate_low = cost_simple_low(a, ctx, b, rate);
    const int64_t rd = rdcost(rdmult, rate, coeff_dist(at, ad, SHIFT));
    const int32_t ad_low = (int32_t)(((int64_t)(a - 1) * dv) >> SHIFT);
    return rdcost(rdmult, rate_low, coeff_dist(at, ad_low, SHIFT)) < rd ? a - 1 : a;
  }
  // level a - 1 at ci in place of a (sign 1 for a negative level)
  __device__ __forceinline__ void store_low(int ci, int a, int sign) {
    const int32_t ad_low = (int32_t)(((int64_t)(a - 1) * dqv[ci != 0]) >> SHIFT);
    qc[ci] = (-sign ^ (a - 1)) + sign;
    dq[ci] = (-sign ^ ad_low) + sign;
  }
  // update_coeff_simple (its running rate is never read again: dropped)
  __device__ void update_simple(int si) {
    int ci;
    int32_t t, qv, d;
    fetch(si, ci, t, qv, d);
    if (qv == 0) return;
    const int a = qv < 0 ? -qv : qv, l = simple_level(ci, si, t, qv);
    if (l != a) {
      store_low(ci, a, qv < 0);
      set_level(ci, l);
    }
  }
  // quant.rs:1120 update_coeff_eob; nz_ci[0 .. nz_num) are the nonzero
  // positions since the eob (at most 5)
  // The step's contexts come from one read of the template levels
  // (ctx_pair) and the eob context and eob cost from the position alone, so
  // its table reads issue together: the sequential walk's chain is scan ->
  // coefficient -> levels -> tables.
  __device__ void update_eob(int* accu_rate, int64_t* accu_dist, int* eob, int* nz_num, int* nz_ci, int si,
                             bool sharp) {
    int ci;
    int32_t t, qv, d;
    fetch(si, ci, t, qv, d);
    const int dv = dqv[ci != 0];
    int ctx, brc;
    ctx_pair(ci, &ctx, &brc);
    if (qv == 0) {
      *accu_rate += base_cost(ctx, 0);
      return;
    }
    bool lower = false;
    const int a = qv < 0 ? -qv : qv, sign = qv < 0;
    const int ctx_ne = si == 0 ? 0 : (si <= N / 8 ? 1 : (si <= N / 4 ? 2 : 3));
    const int brc_ne = br_ctx_eob(ci);
    // every cost term of the four candidates (a, a - 1; before the eob, last)
    const int sgn = ci == 0 ? cost[kDcSign + dc_ctx * 2 + sign] : 512;
    const int b_hi = base_cost(ctx, min(a, 3)), b_lo = base_cost(ctx, min(a - 1, 3));
    const int e_hi = cost[kBaseEob + ctx_ne * 3 + min(a, 3) - 1], e_lo = cost[kBaseEob + ctx_ne * 3 + max(min(a - 1, 3) - 1, 0)];
    const int p_hi = lps(brc, min(max(a - 3, 0), 12)), p_lo = lps(brc, min(max(a - 4, 0), 12));
    const int pe_hi = lps(brc_ne, min(max(a - 3, 0), 12)), pe_lo = lps(brc_ne, min(max(a - 4, 0), 12));
    const int ne_cost = eob_cost(si + 1);
    const int g_hi = golomb_cost(a), g_lo = golomb_cost(a - 1);
    // cost_general(false, ci, a / a - 1) and cost_eob(ci, a / a - 1) (quant.rs:758 / :795)
    const int c_hi = b_hi + sgn + (a > 2 ? p_hi + g_hi : 0);
    const int c_lo = a - 1 == 0 ? b_lo : b_lo + sgn + (a - 1 > 2 ? p_lo + g_lo : 0);
    const int ce_hi = e_hi + sgn + (a > 2 ? pe_hi + g_hi : 0);
    const int ce_lo = e_lo + sgn + (a - 1 > 2 ? pe_lo + g_lo : 0);
    const int64_t dist0 = coeff_dist(t, 0, SHIFT);
    int64_t dist = coeff_dist(t, d, SHIFT) - dist0;
    int rate = c_hi;
    const int64_t rd = rdcost(rdmult, *accu_rate + rate, *accu_dist + dist);
    int32_t q_low = 0, d_low = 0;
    int rate_low;
    int64_t dist_low, rd_low;
    if (a == 1) {
      dist_low = 0;
      rate_low = b_lo;  // base_cost(ctx, 0)
      rd_low = rdcost(rdmult, *accu_rate + rate_low, *accu_dist);
    } else {
      low(a, sign, dv, &q_low, &d_low);
      dist_low = coeff_dist(t, d_low, SHIFT) - dist0;
      rate_low = c_lo;
      rd_low = rdcost(rdmult, *accu_rate + rate_low, *accu_dist + dist_low);
    }
    bool lower_ne = false;
    int rate_ce = ne_cost + ce_hi;
    int64_t dist_ne = dist;
    int64_t rd_ne = rdcost(rdmult, rate_ce, dist_ne);
    if (a - 1 > 0) {
      const int rate_ce_low = ne_cost + ce_lo;
      const int64_t rd_ne_low = rdcost(rdmult, rate_ce_low, dist_low);
      if (rd_ne_low < rd_ne) {
        lower_ne = true;
        rd_ne = rd_ne_low;
        rate_ce = rate_ce_low;
        dist_ne = dist_low;
      }
    }
    if (rd_low < rd) {
      lower = true;
      rate = rate_low;
      dist = dist_low;
    }
    if (!sharp && rd_ne < (rd_low < rd ? rd_low : rd)) {
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k < *nz_num) {
          set_level(nz_ci[k], 0);
          qc[nz_ci[k]] = 0;
          dq[nz_ci[k]] = 0;
        }
      *eob = si + 1;
      *nz_num = 0;
      *accu_rate = rate_ce;
      *accu_dist = dist_ne;
      lower = lower_ne;
    } else {
      *accu_rate += rate;
      *accu_dist += dist;
    }
    if (lower) {
      qc[ci] = q_low;
      dq[ci] = d_low;
      set_level(ci, a - 1);
    }
    if (lower ? q_low != 0 : true) {  // the level kept is nonzero
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k == *nz_num) nz_ci[k] = ci;
      ++*nz_num;
    }
  }
};
}  // namespace

// The part of optimize_b (quant.rs:1312) before the simple walk, for a wave
// that holds one block (every lane runs it with the same values): the last
// coefficient, the eob phase and the skip decision.  Returns false for a
// block the trellis is not defined on (level 0 at scan[eob - 1]: left as it
// is).  On return si is the first scan index of the simple walk and si_end
// its cut-off (rdoq_cutoffs, quant.rs:1551).  A zero level in the eob phase only adds base_cost(ctx, 0)
// to the running rate, with ctx from levels that no zero can change, so the
// zeros down to the next nonzero level are priced by the wave at once (64
// positions a pass, found by ballot) and only nonzero levels take a
// sequential step.
template <int TXS>
__device__ bool trellis_head_wave(Trellis<TXS, WaveMap>& s, int lane, int& eob, int& si, int& si_end,
                                  int non_skip_cost, int skip_cost, bool sharp, uint32_t cut_num, uint32_t cut_den) {
  using T = Trellis<TXS, WaveMap>;
  int accu_rate = s.eob_cost(eob);
  int64_t accu_dist = 0;
  si = eob - 1;
  int ci;
  int32_t t, qv, d;
  s.fetch(si, ci, t, qv, d);
  if (qv == 0) return false;
  {
    const int a = qv < 0 ? -qv : qv, sign = qv < 0;
    if (a >= 2) {
      s.update_general(&accu_rate, &accu_dist, si, eob);
    } else {
      accu_rate += s.cost_eob(ci, a, sign, s.lower_ctx(ci, si, true));
      accu_dist += coeff_dist(t, d, T::SHIFT) - coeff_dist(t, 0, T::SHIFT);
    }
  }
  si--;
  int nz_num = 1, nz_ci[5] = {ci, 0, 0, 0, 0};
  while (si >= 0 && nz_num <= 4) {
    const int i = si - lane;
    int cz = 0;
    bool nz = false;
    if (i >= 0) {
      cz = s.scan[i];
      nz = s.m.qs[cz] != 0;  // below si: still the quantizer's level
    }
    const uint64_t m = __ballot(nz);
    const int nzero = m ? (int)__builtin_ctzll(m) : min(64, si + 1);
    const int r = lane < nzero ? s.base_cost(s.lower_ctx(cz, i, false), 0) : 0;
    accu_rate += group_sum_i32<64>(r);
    si -= nzero;
    if (m == 0) continue;
    s.update_eob(&accu_rate, &accu_dist, &eob, &nz_num, nz_ci, si, sharp);
    si--;
  }
  if (si == -1 && nz_num <= 4) {  // update_skip (quant.rs:1281)
    const int64_t rd = rdcost(s.rdmult, accu_rate + non_skip_cost, accu_dist);
    if (!sharp && rdcost(s.rdmult, skip_cost, 0) < rd) {
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k < nz_num) s.qc[nz_ci[k]] = 0, s.dq[nz_ci[k]] = 0;
      eob = 0;
    }
  }
  si_end = 1;
  if (cut_num != 0) {
    const uint32_t cut = max((uint32_t)(T::N >> 7), (uint32_t)eob * cut_num / cut_den);
    si_end = max(eob - (int)cut, 1);
  }
  return true;
}

template <int TXS>
__device__ __forceinline__ bool job_ok(const RdoqJob& job, int eob) {
  constexpr int N = c_txw(TXS) * c_txh(TXS);
  // out-of-range fields would index past the staged tables: such blocks are left as they are
  return eob > 0 && eob <= N && job.tx_class <= 2 && job.plane_type <= 1 && job.txb_skip_ctx < 13 &&
         job.dc_sign_ctx < 3;
}

template <int TXS>
__global__ __launch_bounds__(64) void k_rdoq(const int2* __restrict__ work, const uint32_t* __restrict__ lvwork,
                                             int32_t* __restrict__ qcoeff, int32_t* __restrict__ dqcoeff,
                                             uint16_t* __restrict__ eobs, const RdoqJob* __restrict__ jobs, int n_jobs,
                                             const int32_t* __restrict__ costs, const uint16_t* __restrict__ scans,
                                             int32_t dq0, int32_t dq1, uint32_t cut_num, uint32_t cut_den) {
  using T = Trellis<TXS, GatherMap>;
  constexpr int N = T::N;
  __shared__ int32_t tab[2 * kTxbWords + 2 * kEobWords];
  __shared__ uint16_t sc[3 * N];
  __shared__ uint32_t lv[T::NW * 64];
  const int lane = threadIdx.x;
  // CoeffCostTables (quant.rs:475): txb[txs_ctx * 2 + plane], then eob[ms][plane]
  const int32_t* txb_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;
  const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
  for (int i = lane; i < 2 * kTxbWords; i += 64) tab[i] = txb_src[i];
  for (int i = lane; i < 2 * kEobWords; i += 64) tab[2 * kTxbWords + i] = eob_src[i];
  for (int i = lane; i < 3 * N; i += 64) sc[i] = scans[i];
  __syncthreads();
  const int j = blockIdx.x * 64 + lane;
  if (j >= n_jobs) return;
  const RdoqJob job = jobs[j];
  int eob = eobs[j];
  if (!job_ok<TXS>(job, eob)) return;
  T s;
  s.m.work = work + (size_t)j * N;
  s.m.lv = lv + lane;
  s.qc = qcoeff + job.coeff;
  s.dq = dqcoeff + job.coeff;
  s.cost = tab + (job.plane_type ? kTxbWords : 0);
  s.ecost = tab + 2 * kTxbWords + (job.plane_type ? kEobWords : 0);
  // scan classes: 0 diagonal (2D), 1 row (vertical class), 2 column (horizontal class)
  s.scan = sc + (job.tx_class == kClsVert ? 1 : (job.tx_class == kClsHoriz ? 2 : 0)) * N;
  s.rdmult = job.rdmult;
  s.dqv[0] = dq0, s.dqv[1] = dq1;
  s.cls = job.tx_class;
  s.dc_ctx = job.dc_sign_ctx;
  const bool sharp = job.sharpness != 0;
  // txb_init_levels (coeff_c.rs:278), packed by the gather
  for (int w = 0; w < T::NW; w++) s.m.lv[w * 64] = lvwork[(long)j * T::NW + w];
  const int non_skip_cost = s.cost[kSkip + job.txb_skip_ctx * 2], skip_cost = s.cost[kSkip + job.txb_skip_ctx * 2 + 1];
  int accu_rate = s.eob_cost(eob);
  int64_t accu_dist = 0;
  int si = eob - 1;
  {
    int ci;
    int32_t t, qv, d;
    s.fetch(si, ci, t, qv, d);
    // an eob that is not tight (level 0 at scan[eob - 1]) is not a block the
    // trellis is defined on: left as it is
    if (qv == 0) return;
    const int a = qv < 0 ? -qv : qv, sign = qv < 0;
    if (a >= 2) {
      s.update_general(&accu_rate, &accu_dist, si, eob);
    } else {
      accu_rate += s.cost_eob(ci, a, sign, s.lower_ctx(ci, si, true));
      accu_dist += coeff_dist(t, d, T::SHIFT) - coeff_dist(t, 0, T::SHIFT);
    }
    si--;
  }
  int nz_num = 1, nz_ci[5] = {s.scan[eob - 1], 0, 0, 0, 0};
  while (si >= 0 && nz_num <= 4) {
    s.update_eob(&accu_rate, &accu_dist, &eob, &nz_num, nz_ci, si, sharp);
    si--;
  }
  if (si == -1 && nz_num <= 4) {  // update_skip (quant.rs:1281)
    const int64_t rd = rdcost(s.rdmult, accu_rate + non_skip_cost, accu_dist);
    if (!sharp && rdcost(s.rdmult, skip_cost, 0) < rd) {
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k < nz_num) s.qc[nz_ci[k]] = 0, s.dq[nz_ci[k]] = 0;
      eob = 0;
    }
  }
  int si_end = 1;
  if (cut_num != 0) {
    const uint32_t cut = max((uint32_t)(N >> 7), (uint32_t)eob * cut_num / cut_den);
    si_end = max(eob - (int)cut, 1);
  }
  for (; si >= si_end; si--) s.update_simple(si);
  if (si == 0) {
    int64_t dummy = 0;
    s.update_general(&accu_rate, &dummy, 0, eob);
  }
  eobs[j] = (uint16_t)eob;
}

// The wave form: kWaveBlocks blocks per workgroup, one wave each.
constexpr int kWaveBlocks = 4;
template <int TXS>
__global__ __launch_bounds__(64 * kWaveBlocks) void k_rdoq_wave(
    const int32_t* __restrict__ tcoeff, int32_t* __restrict__ qcoeff, int32_t* __restrict__ dqcoeff,
    uint16_t* __restrict__ eobs, const RdoqJob* __restrict__ jobs, int n_jobs, const int32_t* __restrict__ costs,
    const uint16_t* __restrict__ scans, int32_t dq0, int32_t dq1, uint32_t cut_num, uint32_t cut_den) {
  using T = Trellis<TXS, WaveMap>;
  constexpr int N = T::N, K = (N + 63) / 64;
  __shared__ int32_t tab[2 * kTxbWords + 2 * kEobWords];
  __shared__ uint16_t sc[3 * N];
  __shared__ int32_t ts[kWaveBlocks][N], qs[kWaveBlocks][N];
  __shared__ uint8_t lvb[kWaveBlocks][N];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int32_t* txb_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;
  const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
  for (int i = tid; i < 2 * kTxbWords; i += 64 * kWaveBlocks) tab[i] = txb_src[i];
  for (int i = tid; i < 2 * kEobWords; i += 64 * kWaveBlocks) tab[2 * kTxbWords + i] = eob_src[i];
  for (int i = tid; i < 3 * N; i += 64 * kWaveBlocks) sc[i] = scans[i];
  __syncthreads();
  const int j = blockIdx.x * kWaveBlocks + wv;
  if (j >= n_jobs) return;  // wave-uniform; no barrier follows
  const RdoqJob job = jobs[j];
  int eob = eobs[j];
  if (!job_ok<TXS>(job, eob)) return;
  // the block, raster, and its levels (txb_init_levels, coeff_c.rs:278)
  for (int p = lane; p < N; p += 64) {
    const int32_t t = tcoeff[job.coeff + p], q = qcoeff[job.coeff + p];
    ts[wv][p] = t;
    qs[wv][p] = q;
    const int a = q < 0 ? -q : q;
    lvb[wv][p] = (uint8_t)(a < 15 ? a : 15);
  }
  wave_sync();
  T s;
  s.m.lv = lvb[wv];
  s.m.ts = ts[wv];
  s.m.qs = qs[wv];
  s.qc = qcoeff + job.coeff;
  s.dq = dqcoeff + job.coeff;
  s.cost = tab + (job.plane_type ? kTxbWords : 0);
  s.ecost = tab + 2 * kTxbWords + (job.plane_type ? kEobWords : 0);
  s.scan = sc + (job.tx_class == kClsVert ? 1 : (job.tx_class == kClsHoriz ? 2 : 0)) * N;
  s.rdmult = job.rdmult;
  s.dqv[0] = dq0, s.dqv[1] = dq1;
  s.cls = job.tx_class;
  s.dc_ctx = job.dc_sign_ctx;
  int si = 0, si_end = 1;
  if (!trellis_head_wave(s, lane, eob, si, si_end, s.cost[kSkip + job.txb_skip_ctx * 2],
                         s.cost[kSkip + job.txb_skip_ctx * 2 + 1], job.sharpness != 0, cut_num, cut_den))
    return;
  wave_sync();
  // the simple walk over scan indices [si_end, si] as a fixpoint; lane l
  // holds positions si - l - 64 k, re-read from LDS each pass (the
  // quantizer's levels there are never changed by the walk), its lowered
  // positions as bits of one word
  static_assert(K <= 32, "one bit per position of a lane");
  if (si >= si_end) {
    uint32_t lowm = 0;
    for (;;) {
      bool changed = false;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = si - lane - 64 * k;
        if (i < si_end) break;
        const int ci = s.scan[i];
        const int32_t qv = s.m.qs[ci];
        if (qv == 0) continue;
        const int a = qv < 0 ? -qv : qv;
        const int l = s.simple_level(ci, i, s.m.ts[ci], qv);
        const bool lo = l != a;
        if (lo != ((lowm >> k) & 1)) {
          lowm ^= 1u << k;
          s.set_level(ci, l);
          changed = true;
        }
      }
      wave_sync();
      if (__ballot(changed) == 0) break;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
      if ((lowm >> k) & 1) {
        const int ci = s.scan[si - lane - 64 * k];
        const int32_t qv = s.m.qs[ci];
        s.store_low(ci, qv < 0 ? -qv : qv, qv < 0);
      }
  }
  // the walk's last step: the DC position when the cut-off reaches it
  const int last = si >= si_end ? si_end - 1 : si;
  if (last == 0) {
    int accu_rate = 0;
    int64_t dummy = 0;
    s.update_general(&accu_rate, &dummy, 0, eob);
  }
  if (lane == 0) eobs[j] = (uint16_t)eob;
}

// The group form: G lanes per block (64 / G blocks per wave, 4 waves per
// workgroup).  The eob phase is sequential per block and runs the same
// instructions on every lane of its group, so a wave advances 64 / G
// blocks' walks at once instead of one; the zeros between nonzero levels
// are priced G at a time, the fixpoint gives each lane N / G positions.
// Everything else is the wave form's.  Collective steps (ballots, group
// sums) run in wave-uniform control flow; a group whose block is done or
// not defined keeps taking part with no effect.
// CMP (compact): every job is luma and every level fits 16 bits (the
// caller's guarantee), so the workgroup stages one plane's cost tables and
// 16-bit levels -- less LDS per block, more blocks resident per CU.
// One block of the group form on the G lanes of its group (lane g of the
// group, group grp of the wave): stage it into the group's LDS slot, the
// eob phase, the skip decision, the fixpoint walk and the DC step.  live:
// j names a job.  Every lane of the wave calls it together (ballots, DPP
// sums, wave syncs); a group without a live, defined job takes part with
// no effect.
template <int TXS, int G, bool CMP, class QT>
__device__ __forceinline__ void group_trellis(int j, bool live, int lane, int g, int grp,
                                              const int32_t* __restrict__ tcoeff, int32_t* __restrict__ qcoeff,
                                              int32_t* __restrict__ dqcoeff, uint16_t* __restrict__ eobs,
                                              const RdoqJob* __restrict__ jobs, const int32_t* tab,
                                              const uint16_t* sc, int32_t* tsb, QT* qsb, uint8_t* lvbb, int32_t dq0,
                                              int32_t dq1, uint32_t cut_num, uint32_t cut_den) {
  using T = Trellis<TXS, WaveMapQ<QT>>;
  constexpr int N = T::N, K = N / G;
  constexpr int TXB = CMP ? kTxbWords : 2 * kTxbWords;
  RdoqJob job{};
  int eob = 0;
  bool act = false;
  if (live) {
    job = jobs[j];
    eob = eobs[j];
    act = job_ok<TXS>(job, eob) && (!CMP || job.plane_type == 0);
  }
  // the block, raster, and its levels (txb_init_levels, coeff_c.rs:278)
  if (act)
    for (int p = g; p < N; p += G) {
      const int32_t t = tcoeff[job.coeff + p], q = qcoeff[job.coeff + p];
      tsb[p] = t;
      qsb[p] = (QT)q;
      const int a = q < 0 ? -q : q;
      lvbb[p] = (uint8_t)(a < 15 ? a : 15);
    }
  wave_sync();
  T s;
  s.m.lv = lvbb;
  s.m.ts = tsb;
  s.m.qs = qsb;
  s.qc = qcoeff + job.coeff;
  s.dq = dqcoeff + job.coeff;
  s.cost = tab + (job.plane_type ? kTxbWords : 0);
  s.ecost = tab + TXB + (job.plane_type ? kEobWords : 0);
  s.scan = sc + (job.tx_class == kClsVert ? 1 : (job.tx_class == kClsHoriz ? 2 : 0)) * N;
  s.rdmult = job.rdmult;
  s.dqv[0] = dq0, s.dqv[1] = dq1;
  s.cls = job.tx_class;
  s.dc_ctx = job.dc_sign_ctx;
  const bool sharp = job.sharpness != 0;
  // ---- the last coefficient (as the wave form's head) ----
  int accu_rate = 0, si = 0, nz_num = 0, nz_ci[5] = {0, 0, 0, 0, 0};
  int64_t accu_dist = 0;
  if (act) {
    accu_rate = s.eob_cost(eob);
    si = eob - 1;
    int ci;
    int32_t t, qv, d;
    s.fetch(si, ci, t, qv, d);
    if (qv == 0) {
      act = false;  // not defined: the block is left as it is
    } else {
      const int a = qv < 0 ? -qv : qv, sign = qv < 0;
      if (a >= 2) {
        s.update_general(&accu_rate, &accu_dist, si, eob);
      } else {
        accu_rate += s.cost_eob(ci, a, sign, s.lower_ctx(ci, si, true));
        accu_dist += coeff_dist(t, d, T::SHIFT) - coeff_dist(t, 0, T::SHIFT);
      }
      si--;
      nz_num = 1, nz_ci[0] = ci;
    }
  }
  // ---- the eob phase, G positions of zeros at a time ----
  for (;;) {
    const bool walking = act && si >= 0 && nz_num <= 4;
    if (__ballot(walking) == 0) break;
    const int i = si - g;
    int cz = 0;
    bool nz = false;
    if (walking && i >= 0) {
      cz = s.scan[i];
      nz = s.m.qs[cz] != 0;  // below si: still the quantizer's level
    }
    const uint32_t m = (uint32_t)(__ballot(nz) >> (grp * G)) & (G == 32 ? 0xffffffffu : ((1u << G) - 1));
    const int nzero = m ? (int)__builtin_ctz(m) : min(G, si + 1);
    const int r = (walking && g < nzero) ? s.base_cost(s.lower_ctx(cz, i, false), 0) : 0;
    const int rs = dpp_sum_i32<G>(r);  // DPP row rotations: no LDS round trip
    if (walking) {
      accu_rate += rs;
      si -= nzero;
      if (m != 0) {
        s.update_eob(&accu_rate, &accu_dist, &eob, &nz_num, nz_ci, si, sharp);
        si--;
      }
    }
  }
  int si_end = 1;
  if (act) {
    if (si == -1 && nz_num <= 4) {  // update_skip (quant.rs:1281)
      const int64_t rd = rdcost(s.rdmult, accu_rate + s.cost[kSkip + job.txb_skip_ctx * 2], accu_dist);
      if (!sharp && rdcost(s.rdmult, s.cost[kSkip + job.txb_skip_ctx * 2 + 1], 0) < rd) {
#pragma unroll
        for (int k = 0; k < 5; k++)
          if (k < nz_num) s.qc[nz_ci[k]] = 0, s.dq[nz_ci[k]] = 0;
        eob = 0;
      }
    }
    if (cut_num != 0) {
      const uint32_t cut = max((uint32_t)(N >> 7), (uint32_t)eob * cut_num / cut_den);
      si_end = max(eob - (int)cut, 1);
    }
  }
  wave_sync();
  // ---- the simple walk over [si_end, si] as a fixpoint (the wave form's),
  // lane g of the group holding positions si - g - G k ----
  const bool fix = act && si >= si_end;
  uint32_t lowm = 0;
  for (;;) {
    bool changed = false;
    if (fix) {
#pragma unroll 4
      for (int k = 0; k < K; k++) {
        const int i = si - g - G * k;
        if (i < si_end) break;
        const int ci = s.scan[i];
        const int32_t qv = s.m.qs[ci];
        if (qv == 0) continue;
        const int a = qv < 0 ? -qv : qv;
        const int l = s.simple_level(ci, i, s.m.ts[ci], qv);
        const bool lo = l != a;
        if (lo != ((lowm >> k) & 1)) {
          lowm ^= 1u << k;
          s.set_level(ci, l);
          changed = true;
        }
      }
    }
    wave_sync();
    if (__ballot(changed) == 0) break;
  }
  if (fix) {
    for (int k = 0; k < K; k++)
      if ((lowm >> k) & 1) {
        const int ci = s.scan[si - g - G * k];
        const int32_t qv = s.m.qs[ci];
        s.store_low(ci, qv < 0 ? -qv : qv, qv < 0);
      }
  }
  if (act) {
    // the walk's last step: the DC position when the cut-off reaches it
    const int last = si >= si_end ? si_end - 1 : si;
    if (last == 0) {
      int ar = 0;
      int64_t dummy = 0;
      s.update_general(&ar, &dummy, 0, eob);
    }
    if (g == 0) eobs[j] = (uint16_t)eob;
  }
}

template <int TXS, int G, int WPG, bool CMP>
__global__ __launch_bounds__(64 * WPG) void k_rdoq_group(
    const int32_t* __restrict__ tcoeff, int32_t* __restrict__ qcoeff, int32_t* __restrict__ dqcoeff,
    uint16_t* __restrict__ eobs, const RdoqJob* __restrict__ jobs, int n_jobs, const int32_t* __restrict__ costs,
    const uint16_t* __restrict__ scans, int32_t dq0, int32_t dq1, uint32_t cut_num, uint32_t cut_den) {
  using QT = typename std::conditional<CMP, int16_t, int32_t>::type;
  constexpr int N = c_txw(TXS) * c_txh(TXS), BPW = 64 / G, BPG = WPG * BPW, K = N / G, NT = 64 * WPG;
  constexpr int TXB = CMP ? kTxbWords : 2 * kTxbWords, EOBW = CMP ? kEobWords : 2 * kEobWords;
  static_assert(N % G == 0 && K <= 32, "positions of a lane as bits of one word");
  static_assert(G <= 16, "group sums by DPP within a row of 16 lanes");
  __shared__ int32_t tab[TXB + EOBW];
  __shared__ uint16_t sc[3 * N];
  __shared__ int32_t ts[BPG][N];
  __shared__ QT qs[BPG][N];
  __shared__ uint8_t lvb[BPG][N];
  const int tid = threadIdx.x, lane = tid & 63, g = lane % G, grp = lane / G, b = tid / G;
  // the size's tables: both planes, or the luma plane's (the first of each pair)
  const int32_t* txb_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;
  const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
  for (int i = tid; i < TXB; i += NT) tab[i] = txb_src[i];
  for (int i = tid; i < EOBW; i += NT) tab[TXB + i] = eob_src[i];
  for (int i = tid; i < 3 * N; i += NT) sc[i] = scans[i];
  __syncthreads();
  const int j = blockIdx.x * BPG + b;
  group_trellis<TXS, G, CMP, QT>(j, j < n_jobs, lane, g, grp, tcoeff, qcoeff, dqcoeff, eobs, jobs, tab, sc, ts[b],
                                 qs[b], lvb[b], dq0, dq1, cut_num, cut_den);
}

// The persistent compact group form fed by a device work queue: queue[0 ..
// ctr[0]) holds the jobs to walk (written by the pre-trellis gate: the jobs
// it passes with a nonzero eob), ctr[1] is the queue head.  A grid of about
// the device's resident capacity stages the cost tables once per workgroup,
// then each wave takes the next 64 / G jobs (one atomic per wave) until the
// queue is empty -- every wave reaches that exit.  Jobs outside the queue
// are left as they are, as the direct form leaves the jobs it skips.
template <int TXS, int G, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_rdoq_queue(
    const int32_t* __restrict__ tcoeff, int32_t* __restrict__ qcoeff, int32_t* __restrict__ dqcoeff,
    uint16_t* __restrict__ eobs, const RdoqJob* __restrict__ jobs, const int32_t* __restrict__ queue,
    unsigned* __restrict__ ctr, const int32_t* __restrict__ costs, const uint16_t* __restrict__ scans, int32_t dq0,
    int32_t dq1, uint32_t cut_num, uint32_t cut_den) {
  using QT = int16_t;
  constexpr int N = c_txw(TXS) * c_txh(TXS), BPW = 64 / G, BPG = WPG * BPW, K = N / G, NT = 64 * WPG;
  static_assert(N % G == 0 && K <= 32 && G <= 16, "group form shape");
  __shared__ int32_t tab[kTxbWords + kEobWords];
  __shared__ uint16_t sc[3 * N];
  __shared__ int32_t ts[BPG][N];
  __shared__ QT qs[BPG][N];
  __shared__ uint8_t lvb[BPG][N];
  const int tid = threadIdx.x, lane = tid & 63, g = lane % G, grp = lane / G, b = tid / G;
  const int32_t* txb_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;
  const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
  for (int i = tid; i < kTxbWords; i += NT) tab[i] = txb_src[i];
  for (int i = tid; i < kEobWords; i += NT) tab[kTxbWords + i] = eob_src[i];
  for (int i = tid; i < 3 * N; i += NT) sc[i] = scans[i];
  __syncthreads();
  const unsigned qlen = ctr[0];
  for (;;) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&ctr[1], (unsigned)BPW);
    base = (unsigned)__shfl((int)base, 0, 64);
    if (base >= qlen) break;  // wave-uniform
    const unsigned qi = base + (unsigned)grp;
    const int j = qi < qlen ? queue[qi] : -1;
    group_trellis<TXS, G, true, QT>(j, j >= 0, lane, g, grp, tcoeff, qcoeff, dqcoeff, eobs, jobs, tab, sc, ts[b],
                                    qs[b], lvb[b], dq0, dq1, cut_num, cut_den);
    wave_sync();  // the group's LDS slot is rewritten by the next job
  }
}

// Coefficient rate (cost_coeffs_txb, lf/coeff_rate.rs:60, full middle loop)
// of quantized blocks.  Without a trellis every scanned position's cost is a
// function of the final level map, so the positions of a block are priced in
// parallel: a group of G = min(16, N) lanes per block up to 256 coefficients
// (16 blocks per workgroup, so the cost tables staged in LDS serve 16
// blocks), 64 above (the blocks' LDS), walks the
// scan indices below the eob with stride G, reading levels straight from the
// block (min(|q|, 127), 0 past its edges), and the group's sum is reduced
// across its lanes; the group's first lane adds the skip-flag, tx-type and
// eob costs.
template <int TXS>
__global__ __launch_bounds__(256) void k_coeff_rate(const int32_t* __restrict__ qcoeff,
                                                    const uint16_t* __restrict__ eobs,
                                                    const RateJob* __restrict__ jobs, int n_jobs,
                                                    const int32_t* __restrict__ costs,
                                                    const uint16_t* __restrict__ scans, int32_t* __restrict__ rate) {
  constexpr int AW = c_txw(TXS), AH = c_txh(TXS), N = AW * AH;
  constexpr int G = N < 16 ? N : (N <= 256 ? 16 : 64), PER_WG = 256 / G;  // LDS: PER_WG blocks
  __shared__ int32_t tab[2 * kTxbWords + 2 * kEobWords];
  __shared__ uint16_t sc[3 * N];
  __shared__ int32_t qs[PER_WG * N];  // each group's block, raster
  const int tid = threadIdx.x;
  const int32_t* txb_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;
  const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
  for (int i = tid; i < 2 * kTxbWords; i += 256) tab[i] = txb_src[i];
  for (int i = tid; i < 2 * kEobWords; i += 256) tab[2 * kTxbWords + i] = eob_src[i];
  for (int i = tid; i < 3 * N; i += 256) sc[i] = scans[i];
  __syncthreads();
  const int g = tid % G, j = blockIdx.x * PER_WG + tid / G;
  const bool live = j < n_jobs;
  RateJob job{};
  int eob = 0;
  if (live) {
    job = jobs[j];
    eob = eobs[j];
    if (eob > N || job.tx_class > 2 || job.plane_type > 1 || job.txb_skip_ctx >= 13 || job.dc_sign_ctx >= 3)
      eob = -1;  // invalid fields: the block is not priced
  }
  // the group's block into LDS, coalesced; the group is one wave's lanes
  int32_t* q = qs + (tid / G) * N;
  if (eob > 0)
    for (int p = g; p < N; p += G) q[p] = qcoeff[job.coeff + p];
  wave_sync();
  const int32_t* cost = tab + (job.plane_type ? kTxbWords : 0);
  const int32_t* ecost = tab + 2 * kTxbWords + (job.plane_type ? kEobWords : 0);
  const uint16_t* scan = sc + (job.tx_class == kClsVert ? 1 : (job.tx_class == kClsHoriz ? 2 : 0)) * N;
  const int cls = job.tx_class;
  auto lv = [q](int r, int c) -> int {
    if (r >= AH || c >= AW) return 0;
    return min(abs(q[r * AW + c]), 127);
  };
  int32_t sum = 0;
  for (int i = g; i < eob; i += G) {
    const int pos = scan[i], v = q[pos], l = abs(v);
    const bool is_eob = i == eob - 1;
    const int ctx = nz_ctx<TXS>(lv, pos, i, is_eob, cls);
    if (is_eob) {
      sum += cost[kBaseEob + ctx * 3 + max(min(l, 3) - 1, 0)];
    } else if (l > 2 && i > 0) {
      sum += cost[kBase + ctx * 8 + 3];
    } else {
      sum += cost[kBase + ctx * 8 + min(l, 3)];
    }
    if (v != 0) {
      sum += pos == 0 ? cost[kDcSign + job.dc_sign_ctx * 2 + (v < 0)] : 512;
      if (l > 2) sum += cost[kLps + br_ctx_of<TXS>(lv, pos, cls) * 26 + min(l - 3, 12)] + golomb_cost(l);
    }
  }
  // group sum: 16-lane rows by DPP, then across rows
  sum = dpp_sum_i32<(G < 16 ? G : 16)>(sum);
  if constexpr (G >= 32) sum += __shfl_xor(sum, 16, 64);
  if constexpr (G >= 64) sum += __shfl_xor(sum, 32, 64);
  if (live && g == 0) {
    int32_t r = 0;
    if (eob > 0) {
      r = sum + cost[kSkip + job.txb_skip_ctx * 2] + (job.plane_type == 0 ? job.txt_rate : 0);
      // get_eob_cost (quant.rs:851)
      int t;
      if (eob < 33) {
        t = eob <= 2 ? eob : (eob <= 4 ? 3 : (eob <= 8 ? 4 : (eob <= 16 ? 5 : 6)));
      } else {
        const int e = min((eob - 1) >> 5, 16);
        t = e == 0 ? 6 : (e == 1 ? 7 : (e <= 3 ? 8 : (e <= 7 ? 9 : (e <= 15 ? 10 : 11))));
      }
      const int start = t == 0 ? 0 : (t == 1 ? 1 : (1 << (t - 2)) + 1);
      r += ecost[(cls != kCls2D) * 11 + t - 1];
      if (t > 2) {
        const int cnt = t - 3, bit = ((eob - start) >> cnt) & 1;
        r += cost[kEobExtra + cnt * 2 + bit] + 512 * cnt;
      }
    }
    rate[j] = r;
  }
}

hipError_t coeff_rate_batch(int tx_size, const int32_t* qcoeff, const uint16_t* eobs, const RateJob* jobs,
                            int n_jobs, const int32_t* costs, const uint16_t* scans, int32_t* rate, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
#define ZC(T)                                                                                                \
  case T: {                                                                                                  \
    constexpr int N = c_txw(T) * c_txh(T), PER_WG = 256 / (N < 16 ? N : (N <= 256 ? 16 : 64));              \
    hipLaunchKernelGGL((k_coeff_rate<T>), dim3((unsigned)((n_jobs + PER_WG - 1) / PER_WG)), dim3(256), 0, st, \
                       qcoeff, eobs, jobs, n_jobs, costs, scans, rate);                                      \
    break;                                                                                                   \
  }
  switch (tx_size) {
    ZC(0) ZC(1) ZC(2) ZC(3) ZC(4) ZC(5) ZC(6) ZC(7) ZC(8) ZC(9) ZC(10) ZC(11) ZC(12) ZC(13) ZC(14) ZC(15) ZC(16)
    ZC(17) ZC(18)
    default:
      return hipErrorInvalidValue;
  }
#undef ZC
  return hipGetLastError();
}

// (t, q) of every scan index of every block, in scan order: the trellis then
// reads consecutive pairs as its scan index walks down, instead of one
// cache line of a raster row per step.  A workgroup reads whole blocks in
// raster order (coalesced) into LDS and writes them out permuted.
template <int TXS>
__global__ __launch_bounds__(256) void k_rdoq_gather(const int32_t* __restrict__ tcoeff,
                                                     const int32_t* __restrict__ qcoeff,
                                                     const uint16_t* __restrict__ eobs,
                                                     const RdoqJob* __restrict__ jobs, int n_jobs,
                                                     const uint16_t* __restrict__ scans, int2* __restrict__ work,
                                                     uint32_t* __restrict__ lvwork) {
  // 1024 coefficients per workgroup (four per thread): enough loads in
  // flight per workgroup to hide their latency
  constexpr int N = c_txw(TXS) * c_txh(TXS), BPW = N < 1024 ? 1024 / N : 1;
  __shared__ uint16_t sc[3 * N];
  __shared__ int32_t st[BPW * N], sq[BPW * N];
  __shared__ uint32_t boff[BPW];  // per block: coefficient offset, or ~0 when it is not read
  __shared__ int bscan[BPW];      // per block: its scan's offset in sc
  const int tid = threadIdx.x, j0 = blockIdx.x * BPW;
  for (int i = tid; i < 3 * N; i += 256) sc[i] = scans[i];
  if (tid < BPW) {
    const int j = j0 + tid;
    // a block the trellis leaves alone (eob 0) is not read
    const bool use = j < n_jobs && eobs[j] != 0;
    const RdoqJob job = use ? jobs[j] : RdoqJob{};
    const int cls = job.tx_class > 2 ? 0 : job.tx_class;
    boff[tid] = use ? job.coeff : ~0u;
    bscan[tid] = (cls == kClsVert ? 1 : (cls == kClsHoriz ? 2 : 0)) * N;
  }
  __syncthreads();
  for (int k = tid; k < BPW * N; k += 256) {
    const int b = k / N, p = k - b * N;
    const uint32_t off = boff[b];
    if (off != ~0u) {
      st[k] = tcoeff[off + p];
      sq[k] = qcoeff[off + p];
    }
  }
  __syncthreads();
  for (int k = tid; k < BPW * N; k += 256) {
    const int b = k / N, i = k - b * N;
    if (boff[b] != ~0u) {
      const int ci = sc[bscan[b] + i];
      work[(long)(j0 + b) * N + i] = make_int2(st[b * N + ci], sq[b * N + ci]);
    }
  }
  // the blocks' level maps, as the trellis keeps them: min(|q|, 15) in
  // nibbles, eight raster positions per word
  constexpr int NW = (N + 7) / 8;
  for (int k = tid; k < BPW * NW; k += 256) {
    const int b = k / NW, w = k - b * NW;
    if (boff[b] != ~0u) {
      uint32_t v = 0;
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (8 * w + e < N) v |= (uint32_t)min(abs(sq[b * N + 8 * w + e]), 15) << (4 * e);
      lvwork[(long)(j0 + b) * NW + w] = v;
    }
  }
}

// Smallest batch per tx size that takes the lane form when a workspace is
// given.  Default 131 072: measured on MI355X the wave form is ahead inside
// the 4096x4096 wavefront up to its largest steps (~41 000 mostly sparse
// blocks at 64 images: 1.05 vs 1.66 s per pass), the lane form at 262 144
// dense 16x16 blocks (1.49 vs 2.07 ms).
static int g_lane_min[19] = {131072, 131072, 131072, 131072, 131072, 131072, 131072, 131072, 131072, 131072,
                             131072, 131072, 131072, 131072, 131072, 131072, 131072, 131072, 131072};

int rdoq_lane_min(int tx_size) { return tx_size >= 0 && tx_size < 19 ? g_lane_min[tx_size] : 0; }
void rdoq_set_lane_min(int tx_size, int n) {
  if (tx_size >= 0 && tx_size < 19) g_lane_min[tx_size] = n < 0 ? 0 : n;
}

size_t rdoq_workspace(int tx_size, int n_jobs) {
  if (tx_size < 0 || tx_size > 18 || n_jobs <= 0) return 0;
  const int n = std::min(tx_w(tx_size), 32) * std::min(tx_h(tx_size), 32);
  const int nw = (n + 7) / 8;
  return (size_t)n_jobs * (n * sizeof(int2) + nw * sizeof(uint32_t));
}

// The latency form of a size: the group form up to 256 coefficients (16
// lanes per block), the wave form above (a workgroup of group-form blocks
// would not fit its LDS).
template <int TXS>
static void launch_latency_form(const int32_t* tcoeff, int32_t* qcoeff, int32_t* dqcoeff, uint16_t* eobs,
                                const RdoqJob* jobs, int n_jobs, const int32_t* costs, const uint16_t* scans,
                                int32_t dq0, int32_t dq1, uint32_t cut_num, uint32_t cut_den, hipStream_t st,
                                bool compact, const int32_t* queue, unsigned* qctr) {
  constexpr int N = c_txw(TXS) * c_txh(TXS);
  if constexpr (N <= 256) {
    // 12 blocks per workgroup in the compact form (26.9 KB of LDS at 16x16:
    // six workgroups per CU), 16 otherwise (46.1 KB: three)
    if (compact && queue) {
      // persistent: about the resident capacity (six workgroups per CU),
      // never more workgroups than the jobs could fill
      constexpr int BPG = 3 * (64 / 16);
      int dev = 0, cus = 256;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
      const int grid = std::max(1, std::min((n_jobs + BPG - 1) / BPG, cus * 6));
      hipLaunchKernelGGL((k_rdoq_queue<TXS, 16, 3>), dim3((unsigned)grid), dim3(192), 0, st, tcoeff, qcoeff, dqcoeff,
                         eobs, jobs, queue, qctr, costs, scans, dq0, dq1, cut_num, cut_den);
    } else if (compact) {
      constexpr int BPG = 3 * (64 / 16);
      hipLaunchKernelGGL((k_rdoq_group<TXS, 16, 3, true>), dim3((unsigned)((n_jobs + BPG - 1) / BPG)), dim3(192), 0,
                         st, tcoeff, qcoeff, dqcoeff, eobs, jobs, n_jobs, costs, scans, dq0, dq1, cut_num, cut_den);
    } else {
      constexpr int BPG = 4 * (64 / 16);
      hipLaunchKernelGGL((k_rdoq_group<TXS, 16, 4, false>), dim3((unsigned)((n_jobs + BPG - 1) / BPG)), dim3(256), 0,
                         st, tcoeff, qcoeff, dqcoeff, eobs, jobs, n_jobs, costs, scans, dq0, dq1, cut_num, cut_den);
    }
  } else {
    hipLaunchKernelGGL((k_rdoq_wave<TXS>), dim3((unsigned)((n_jobs + kWaveBlocks - 1) / kWaveBlocks)),
                       dim3(64 * kWaveBlocks), 0, st, tcoeff, qcoeff, dqcoeff, eobs, jobs, n_jobs, costs, scans, dq0,
                       dq1, cut_num, cut_den);
  }
}

hipError_t rdoq_batch(int tx_size, const int32_t* tcoeff, int32_t* qcoeff, int32_t* dqcoeff, uint16_t* eobs,
                      const RdoqJob* jobs, int n_jobs, const int32_t* costs, const uint16_t* scans, int32_t dq0,
                      int32_t dq1, uint32_t cut_num, uint32_t cut_den, hipStream_t st, void* ws, size_t ws_bytes,
                      bool compact, const int32_t* queue, unsigned* qctr, int form) {
  if (n_jobs <= 0) return hipSuccess;
  if (tx_size < 0 || tx_size > 18) return hipErrorInvalidValue;
  const bool ws_ok = ws && ws_bytes >= rdoq_workspace(tx_size, n_jobs);
  const bool lane = form == 2 ? ws_ok : (form < 0 && ws_ok && n_jobs >= g_lane_min[tx_size]);
  if (form == 1) queue = nullptr;  // the direct latency form
  const int n = std::min(tx_w(tx_size), 32) * std::min(tx_h(tx_size), 32);
  int2* work = static_cast<int2*>(ws);
  uint32_t* lvwork = lane ? reinterpret_cast<uint32_t*>(work + (size_t)n_jobs * n) : nullptr;
  const int bpw = n < 1024 ? 1024 / n : 1;
  const dim3 grid((unsigned)((n_jobs + 63) / 64)), ggrid((unsigned)((n_jobs + bpw - 1) / bpw));
#define ZR(T)                                                                                                      \
  case T:                                                                                                          \
    if (lane) {                                                                                                    \
      hipLaunchKernelGGL((k_rdoq_gather<T>), ggrid, dim3(256), 0, st, tcoeff, qcoeff, eobs, jobs, n_jobs, scans,   \
                         work, lvwork);                                                                           \
      hipLaunchKernelGGL((k_rdoq<T>), grid, dim3(64), 0, st, work, lvwork, qcoeff, dqcoeff, eobs, jobs, n_jobs,    \
                         costs, scans, dq0, dq1, cut_num, cut_den);                                               \
    } else {                                                                                                       \
      launch_latency_form<T>(tcoeff, qcoeff, dqcoeff, eobs, jobs, n_jobs, costs, scans, dq0, dq1, cut_num, cut_den, \
                             st, compact, queue, qctr);                                                            \
    }                                                                                                              \
    break;
  switch (tx_size) {
    ZR(0) ZR(1) ZR(2) ZR(3) ZR(4) ZR(5) ZR(6) ZR(7) ZR(8) ZR(9) ZR(10) ZR(11) ZR(12) ZR(13) ZR(14) ZR(15) ZR(16)
    ZR(17) ZR(18)
    default:
      return hipErrorInvalidValue;
  }
#undef ZR
  return hipGetLastError();
}

}  // namespace zav1
