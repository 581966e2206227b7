// Fused transform-unit evaluation for the tx-type search, gfx950.
//
// One launch evaluates, for every block of one transform size and every tx
// type of a candidate list: DC prediction from the source neighbours ->
// residual -> forward 2D transform -> 64-point fold with the energy of the
// discarded coefficients -> fp quantization -> frequency-domain distortion,
// eob (scan order) and the level sum.  Nothing but the three per-candidate
// results leaves the chip: residual, coefficients and levels stay in VGPRs /
// LDS.
//
// Semantics follow rust/crates/svtav1-encoder/src/leaf_funnel/tx_pipeline.rs:268
// (tx_unit_inner: :316 residual, :326 fwd, :335-365 fold, :382 quantize_fp,
// :567-577 distortion), rust/crates/svtav1-encoder/src/quant.rs:318 (eob) and
// rust/crates/svtav1-dsp/src/intra_pred.rs:18 (DC prediction).
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {


// Fused search, one lane group per block (all candidate types of the block):
//   DC prediction from the source neighbours (group reduction), then the
//   distinct column transforms of the type list (at most 4: DCT, ADST,
//   FLIPADST, identity) into LDS, items (variant, column) spread over the
//   group's lanes; then the row pass of every type, items (type, row) spread
//   over the P lanes (P / H types at once), each lane quantizing and measuring
//   its row; per-type group reductions give dist / eob / level sum.
// Types are processed in the plan's order (grouped by row kind, so the T/H
// types of one round share their row transform); outputs go to the caller's
// order via plan.out_k.
template <int W, int H, int P>
struct BlkGeom {
  static constexpr int T = P;        // lanes per block (>= max(W, H), divides 64)
  static constexpr int LS = W + 1;   // padded LDS row: conflict-free column writes
  static constexpr int VS = H * LS;  // one column variant of one block
  static constexpr int PW = W < 32 ? W : 32;
  static constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
};

template <int W, int H, int P, int BD, int NQ>
__global__ __launch_bounds__(256) void k_txsearch(const uint16_t* __restrict__ plane, int stride, int bd,
                                                  const int32_t* __restrict__ xy, int n_blocks, TxPlan plan,
                                                  QuantFpPair qp,
                                                  const uint16_t* __restrict__ iscan_all, int tx_size,
                                                  uint64_t* __restrict__ out_dist,
                                                  int32_t* __restrict__ out_eob,
                                                  uint32_t* __restrict__ out_lsum) {
  using G = BlkGeom<W, H, P>;
  static_assert(P >= W && P >= H && 64 % P == 0, "lanes per block");
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  // input bounds of the 1D passes for residuals of <= 10-bit planes: the
  // passes whose bound allows it run with 24-bit-operand butterflies
  constexpr long long kBCol = fwd_col_bound(SH.s0);
  constexpr long long kBRow = fwd_row_bound(H, CBC, SH.s1, kBCol);
  constexpr int LSC = tx_log_scale(W, H);
  extern __shared__ int32_t lds[];

  const int g = threadIdx.x / G::T, t = threadIdx.x - g * G::T;
  const int nvar = plan.n_var, ntyp = plan.n_types;
  const long b = (long)xcd_chunked(blockIdx.x, gridDim.x) * (blockDim.x / G::T) + g;
  const bool active = b < n_blocks;
  const int bi = active ? (int)b : 0;
  const int x0 = xy[2 * bi], y0 = xy[2 * bi + 1];

  // ---- DC prediction from the source neighbours ----
  uint32_t s_above = 0, s_left = 0;
  if (active && y0 > 0 && t < W) s_above = plane[(long)(y0 - 1) * stride + x0 + t];
  if (active && x0 > 0 && t < H) s_left = plane[(long)(y0 + t) * stride + x0 - 1];
  s_above = (uint32_t)group_sum_i32<G::T>((int32_t)s_above);
  s_left = (uint32_t)group_sum_i32<G::T>((int32_t)s_left);
  int32_t dc;
  if (y0 > 0 && x0 > 0) dc = (int32_t)((s_above + s_left + (W + H) / 2) / (W + H));
  else if (y0 > 0) dc = (int32_t)((s_above + W / 2) / W);
  else if (x0 > 0) dc = (int32_t)((s_left + H / 2) / H);
  else dc = 1 << (bd - 1);

  // ---- column variants -> LDS ----
  // Loads are unconditional (an inactive group reads block 0 and its
  // results are never written out), so they issue back to back.
  // Sample offsets are 32-bit from the plane base (planes hold < 2^31
  // samples), so each load needs no 64-bit address arithmetic.  Column
  // outputs are bounded (see fwd_row_bound), so the rounding shift is 32-bit.
  int32_t* L = lds + g * nvar * G::VS;
  const uint32_t o0 = (uint32_t)y0 * (uint32_t)stride + (uint32_t)x0;
  for (int j = t; j < nvar * W; j += G::T) {
    const int v = j / W, c = j - v * W;
    const int ud = plan.var_ud[v];
    const uint32_t oc = o0 + (uint32_t)c + (ud ? (uint32_t)(H - 1) * (uint32_t)stride : 0u);
    const int32_t step = ud ? -stride : stride;
    int32_t col[H];
#pragma unroll
    for (int r = 0; r < H; r++) col[r] = ((int32_t)plane[oc + (uint32_t)(r * step)] - dc) << SH.s0;
    fwd1d<H, CBC, kBCol>(col, plan.var_kind[v]);
#pragma unroll
    for (int r = 0; r < H; r++) L[v * G::VS + r * G::LS + c] = rshift_round32(col[r], SH.s1);
  }
  __syncthreads();

  // ---- row passes: items (type, row), every quantizer of the launch ----
  constexpr int TPR = G::T / H;
  const int sub = t / H, r = t - sub * H;
  for (int base = 0; base < ntyp; base += TPR) {
    const int ti = base + sub;
    const bool live = active && ti < ntyp;
    uint64_t dist[NQ], fold = 0;
    int32_t eob[NQ], lsum[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) dist[q] = 0, eob[q] = 0, lsum[q] = 0;
    if (live) {
      const int tx_type = plan.type[ti];
      const TypeCfg cfg = type_cfg(tx_type);
      // scan positions of the row's kept coefficients, loaded ahead of the
      // transform so their latency overlaps it (rows >= 32 keep none)
      const int cls = (tx_type >= 10) ? (((tx_type - 10) & 1) ? 2 : 1) : 0;
      const uint16_t* iscan = iscan_all + (tx_size * 3 + cls) * 1024 + (r < 32 ? r : 0) * G::PW;
      constexpr int KW = W < 32 ? W : 32;
      int32_t sp[KW];
#pragma unroll
      for (int c = 0; c < KW; c++) sp[c] = (int32_t)iscan[c] + 1;
      const int32_t* src = L + plan.var_of[ti] * G::VS + r * G::LS;
      int32_t v[W];
#pragma unroll
      for (int c = 0; c < W; c++) v[c] = cfg.lr ? src[W - 1 - c] : src[c];
      fwd1d<W, CBR, kBRow>(v, cfg.row);
      // squared errors: each < 2^30 (|e| < 2^15, see QuantFp), so four of
      // them sum in 32 bits before one 64-bit add
      uint32_t sq4[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) sq4[q] = 0;
#pragma unroll
      for (int c = 0; c < W; c++) {
        int32_t x = rshift_round32(v[c], SH.s2);
        // 2:1 rectangles: x NewSqrt2 / 2^12; |x| stays below 2^18 (the pinned
        // coefficient bound before this scaling), so the product fits 32 bits
        if constexpr (G::RECT2) x = (__mul24(x, 5793) + 2048) >> 12;
        if (r >= 32 || c >= 32) {
          // 64-point fold: discarded energy (|x| < 2^18: the 48-bit square
          // from the 24-bit multiplier's low and high words)
          fold += ((uint64_t)(uint32_t)mul_hi_i24(x, x) << 32) | (uint32_t)__mul24(x, x);
          continue;
        }
        const int32_t a = x < 0 ? -x : x;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          int32_t adq;
          const int32_t l = quant_fp_fast<LSC>(a, (r | c) != 0, qp.q[q], BD, &adq);
          lsum[q] += l;
          eob[q] = max(eob[q], l ? sp[c < KW ? c : 0] : 0);
          // x - dq = sign(x) * (a - |dq|): the squared error needs magnitudes only
          const int32_t e = a - adq;
          sq4[q] += (uint32_t)__mul24(e, e);
        }
        if ((c & 3) == 3 || c == W - 1 || c == 31) {
#pragma unroll
          for (int q = 0; q < NQ; q++) dist[q] += sq4[q], sq4[q] = 0;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const uint64_t dq = group_sum_u64<H>(dist[q] + fold);
      const int32_t eq = group_max_i32<H>(eob[q]);
      const int32_t lq = group_sum_i32<H>(lsum[q]);
      if (live && r == 0) {
        constexpr int SHF = (1 - LSC) * 2;
        const long o = (long)q * n_blocks * ntyp + (long)bi * ntyp + plan.out_k[ti];
        out_dist[o] = SHF < 0 ? (dq << (-SHF)) : (dq >> SHF);
        out_eob[o] = eq;
        out_lsum[o] = (uint32_t)lq;
      }
    }
  }
}

// Lane-per-block form for small blocks: one lane holds a whole block and
// evaluates every candidate type of it in registers.  All lanes of a wave
// take the same column variant and the same type at the same time, so the
// kinds, the scan class and the quantizer parameters are wave-uniform
// (scalar registers, scalar branches), each transform is specialised on its
// kind at compile time and no lane exchanges data: no LDS, no barriers, no
// cross-lane reductions.  Types are processed grouped by column variant
// (plan.var_start), so only the current variant's column outputs are live.
// Scan position + 1 of every coefficient of a block of <= 64 samples, per
// scan class (0 default, 1 raster for V_*, 2 column-major for H_*).
struct LaneScan {
  uint32_t pos[3][16];  // four positions per dword, raster order
};

// Passes the inputs of one arm of a wave-uniform selection through an
// empty volatile asm.  The arms' transforms depend only on values that are
// invariant in the variant / type loops, so without this the compiler
// hoists every arm out of the loop (all four kinds computed up front:
// registers for four, and the 24-bit multiply forms lost).  No instruction
// is emitted.
template <int N>
__device__ __forceinline__ void opaque(int32_t (&x)[N]) {
  static_assert(N % 4 == 0, "groups of four");
#pragma unroll
  for (int i = 0; i < N; i += 4) asm volatile("" : "+v"(x[i]), "+v"(x[i + 1]), "+v"(x[i + 2]), "+v"(x[i + 3]));
}

// Column pass of one variant from the block's residual, held packed as
// 16-bit pairs ((sample - dc) << s0 fits 16 bits for <= 10-bit content).
// Where the type loop below keeps its inputs loop-carried (LOOPC, the
// 64-sample sizes: measured faster there, slower for the smaller ones) the
// arms only need to stay branches; otherwise each arm's inputs pass through
// opaque().
template <int W, int H>
constexpr bool ts_lane_loopc() { return W * H == 64; }

template <int W, int H, int CK>
__device__ __forceinline__ void ts_lane_cols(const uint32_t (&rp)[H][W / 2], int32_t (&co)[H][W]) {
  constexpr int CBC = fwd_cb_col(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr long long kBCol = fwd_col_bound(SH.s0);
  if constexpr (ts_lane_loopc<W, H>()) asm volatile("");  // no if-conversion of the four arms
#pragma unroll
  for (int k = 0; k < W / 2; k++) {
    int32_t wv[H];
#pragma unroll
    for (int r = 0; r < H; r++) wv[r] = (int32_t)rp[r][k];
    if constexpr (!ts_lane_loopc<W, H>()) opaque(wv);
#pragma unroll
    for (int half = 0; half < 2; half++) {
      int32_t x[H];
#pragma unroll
      for (int r = 0; r < H; r++) {
        const int32_t w = wv[CK == 2 ? H - 1 - r : r];  // FLIPADST: up-down flip
        x[r] = half ? (w >> 16) : (int32_t)(int16_t)(w & 0xffff);
      }
      fwd1d<H, CBC, kBCol>(x, CK);
#pragma unroll
      for (int r = 0; r < H; r++) co[r][2 * k + half] = rshift_round32(x[r], SH.s1);
    }
  }
}

template <int W, int H, int RK, int BD, bool Q24, int NQ>
__device__ __forceinline__ void ts_lane_rows(const int32_t (&co)[H][W], const uint32_t* __restrict__ pos,
                                             const QuantFpSet& qs, uint64_t (&dist_out)[NQ], int32_t (&eob_out)[NQ],
                                             uint32_t (&lsum_out)[NQ]) {
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr long long kBCol = fwd_col_bound(SH.s0);
  constexpr long long kBRow = fwd_row_bound(H, CBC, SH.s1, kBCol);
  constexpr int LSC = tx_log_scale(W, H);
  constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
  uint64_t dist[NQ];
  int32_t eob[NQ], lsum[NQ];
  uint32_t sq4[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) dist[q] = 0, eob[q] = 0, lsum[q] = 0, sq4[q] = 0;
  // the type's scan positions, read before the transform and held in scalar
  // registers: a select on a freshly loaded value would otherwise become a
  // branch around the load, per coefficient
  uint32_t pw[W * H / 4];
#pragma unroll
  for (int k = 0; k < W * H / 4; k++) {
    pw[k] = pos[k];
    asm volatile("" : "+s"(pw[k]));
  }
#pragma unroll
  for (int r = 0; r < H; r++) {
    int32_t v[W];
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = RK == 2 ? co[r][W - 1 - c] : co[r][c];  // FLIPADST: left-right flip
    if constexpr (!ts_lane_loopc<W, H>()) opaque(v);
    fwd1d<W, CBR, kBRow>(v, RK);
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = rshift_round32(v[c], SH.s2);
      // 2:1 rectangles: x NewSqrt2 / 2^12 (|x| < 2^18, the product fits 32 bits)
      if constexpr (RECT2) x = (__mul24(x, 5793) + 2048) >> 12;
      const int32_t a = x < 0 ? -x : x;
      // scan position + 1 of this coefficient (<= 64, scalar); min(level << 7, p)
      // is p for a nonzero level, else 0
      const int32_t p = (int32_t)((pw[(r * W + c) >> 2] >> (8 * ((r * W + c) & 3))) & 0xffu);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        int32_t e;  // a - |dequantized|, |e| < 2^15: four squares sum in 32 bits
        const int32_t l = quant_fpl<BD, Q24, true>(a, (r | c) != 0, qs.q[q], &e);
        lsum[q] += l;
        eob[q] = max(eob[q], min(l << 7, p));
        sq4[q] = (uint32_t)mad_i24(e, e, (int32_t)sq4[q]);
      }
      if (((r * W + c) & 3) == 3) {
#pragma unroll
        for (int q = 0; q < NQ; q++) dist[q] += sq4[q], sq4[q] = 0;
      }
    }
  }
  constexpr int SHF = (1 - LSC) * 2;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    dist_out[q] = SHF < 0 ? (dist[q] << (-SHF)) : (dist[q] >> SHF);
    eob_out[q] = eob[q];
    lsum_out[q] = (uint32_t)lsum[q];
  }
}

constexpr int kLaneBlocks = 64;  // blocks per workgroup: one per lane of each wave
constexpr int kStageRow = 17;    // staging row: 16 types + 1 (spreads the blocks' rows over the banks)

// Workgroup of 4 waves over the same 64 blocks; wave v takes column variant
// v (all of its types) for every block, lane l block l.  Results meet in a
// per-workgroup staging block and leave as contiguous rows.
template <int W, int H, int BD, bool Q24, int NQ>
__global__ __launch_bounds__(256) void k_txsearch_lane(const uint16_t* __restrict__ plane, int stride, int bd,
                                                       const int32_t* __restrict__ xy, int n_blocks, TxPlan plan,
                                                       QuantFpSet qs, LaneScan scan,
                                                       uint64_t* __restrict__ out_dist,
                                                       int32_t* __restrict__ out_eob,
                                                       uint32_t* __restrict__ out_lsum) {
  static_assert(W * H % 4 == 0 && W <= 16 && H <= 16, "lane-per-block sizes");
  __shared__ uint64_t st_d[NQ][kLaneBlocks * kStageRow];
  __shared__ int32_t st_e[NQ][kLaneBlocks * kStageRow];
  __shared__ uint32_t st_l[NQ][kLaneBlocks * kStageRow];
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr int LSC = tx_log_scale(W, H);
  const int lane = threadIdx.x & 63;
  const int v = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // this wave's variant
  const long b0 = (long)xcd_chunked(blockIdx.x, gridDim.x) * kLaneBlocks;
  const int ntyp = plan.n_types;

  if (v < plan.n_var) {
    const long b = b0 + lane;
    const int bi = b < n_blocks ? (int)b : 0;  // a lane past the end computes block 0, never copied out
    const int x0 = xy[2 * bi], y0 = xy[2 * bi + 1];
    const uint32_t o0 = (uint32_t)y0 * (uint32_t)stride + (uint32_t)x0;
    const PlaneBuf pb(plane);
    const uint32_t rb = 2u * (uint32_t)stride;  // row pitch in bytes

    // DC prediction from the source neighbours
    uint32_t sa = 0, sl = 0;
    if (y0 > 0) {
#pragma unroll
      for (int c = 0; c < W; c++) sa += (uint32_t)pb.at(2u * (o0 - (uint32_t)stride) + 2u * c, 0);
    }
    if (x0 > 0) {
#pragma unroll
      for (int r = 0; r < H; r++) sl += (uint32_t)pb.at(2u * (o0 - 1u), rb * r);
    }
    int32_t dc;
    if (y0 > 0 && x0 > 0) dc = (int32_t)((sa + sl + (W + H) / 2) / (W + H));
    else if (y0 > 0) dc = (int32_t)((sa + W / 2) / W);
    else if (x0 > 0) dc = (int32_t)((sl + H / 2) / H);
    else dc = 1 << (BD - 1);

    // residual, packed: (sample - dc) << s0 as 16-bit pairs
    uint32_t rp[H][W / 2];
#pragma unroll
    for (int r = 0; r < H; r++)
#pragma unroll
      for (int k = 0; k < W / 2; k++) {
        const int32_t lo = (pb.at(2u * o0 + 4u * k, rb * r) - dc) << SH.s0;
        const int32_t hi = (pb.at(2u * o0 + 4u * k + 2u, rb * r) - dc) << SH.s0;
        rp[r][k] = ((uint32_t)lo & 0xffffu) | ((uint32_t)hi << 16);
      }

    const uint32_t vi = plan.lane_var[v];
    const int ck = (int)(vi & 0xff);
    int32_t co[H][W];
    if (ck == 0) ts_lane_cols<W, H, 0>(rp, co);
    else if (ck == 1) ts_lane_cols<W, H, 1>(rp, co);
    else if (ck == 2) ts_lane_cols<W, H, 2>(rp, co);
    else ts_lane_cols<W, H, 3>(rp, co);
    const int i_end = (int)(vi >> 16);
    for (int i = (int)((vi >> 8) & 0xff); i < i_end; i++) {
      const uint32_t ti = plan.lane_type[i];
      const uint32_t* pos = scan.pos[(ti >> 16) & 0xff];
      uint64_t d[NQ];
      int32_t e[NQ];
      uint32_t l[NQ];
      const int rk = (int)((ti >> 8) & 0xff);
      if (rk == 0) ts_lane_rows<W, H, 0, BD, Q24, NQ>(co, pos, qs, d, e, l);
      else if (rk == 1) ts_lane_rows<W, H, 1, BD, Q24, NQ>(co, pos, qs, d, e, l);
      else if (rk == 2) ts_lane_rows<W, H, 2, BD, Q24, NQ>(co, pos, qs, d, e, l);
      else ts_lane_rows<W, H, 3, BD, Q24, NQ>(co, pos, qs, d, e, l);
      const int k = (int)(ti & 0xff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        st_d[q][lane * kStageRow + k] = d[q];
        st_e[q][lane * kStageRow + k] = e[q];
        st_l[q][lane * kStageRow + k] = l[q];
      }
      if constexpr (ts_lane_loopc<W, H>()) {
        // the column outputs become loop-carried: the row arms stay in the
        // type loop and read co in place
#pragma unroll
        for (int r = 0; r < H; r++) opaque(co[r]);
      }
    }
  }
  __syncthreads();

  // The 64 blocks own one contiguous range of each output array (per
  // quantizer): copy it out from the staging rows, consecutive threads on
  // consecutive entries (entry j -> block j / ntyp by a multiply-shift,
  // exact for j < 1024, ntyp <= 16).
  const long nb = n_blocks - b0 < kLaneBlocks ? n_blocks - b0 : kLaneBlocks;
  const uint32_t mg = (65536u + (uint32_t)ntyp - 1u) / (uint32_t)ntyp;
  for (int j = threadIdx.x; j < nb * ntyp; j += blockDim.x) {
    const int blk = (int)(((uint32_t)j * mg) >> 16), k = j - blk * ntyp;
    const int src = blk * kStageRow + k;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const long o = (long)q * n_blocks * ntyp + b0 * ntyp + j;
      out_dist[o] = st_d[q][src];
      out_eob[o] = st_e[q][src];
      out_lsum[o] = st_l[q][src];
    }
  }
}

// Row form for the 16-point sizes (16x16, 8x16, 16x8, 16x4): one lane per
// row, 64 / H blocks per wave, and every lane of a wave on the same column
// variant / the same type at a time (kinds, scan class and type data are
// wave-uniform and scalar).  Per variant: the lanes' columns (residual held
// packed from the start) are transformed into a per-wave LDS tile; per type
// each lane transforms and quantizes its row with its scan positions
// (preloaded bytes for the three classes), and the block's H lanes reduce.
// A block lives in one wave, so only wave-level ordering is needed.
template <int W, int H, int RK, int BD, bool Q24, int NQ>
__device__ __forceinline__ void ts_row_arm(const int32_t* __restrict__ src, const uint32_t (&pc)[W / 4], int r,
                                           const QuantFpSet& qs, uint64_t (&dist)[NQ], int32_t (&m)[NQ],
                                           int32_t (&lsum)[NQ]) {
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr long long kBCol = fwd_col_bound(SH.s0);
  constexpr long long kBRow = fwd_row_bound(H, CBC, SH.s1, kBCol);
  constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
  int32_t v[W];
#pragma unroll
  for (int c = 0; c < W; c++) v[c] = src[RK == 2 ? W - 1 - c : c];  // FLIPADST: left-right flip
  opaque(v);
  fwd1d<W, CBR, kBRow>(v, RK);
  uint32_t sq4[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) sq4[q] = 0;
#pragma unroll
  for (int c = 0; c < W; c++) {
    int32_t x = rshift_round32(v[c], SH.s2);
    if constexpr (RECT2) x = (__mul24(x, 5793) + 2048) >> 12;  // |x| < 2^18
    const int32_t a = x < 0 ? -x : x;
    const int32_t p = (int32_t)((pc[c >> 2] >> (8 * (c & 3))) & 0xffu);  // scan index
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      int32_t e;
      // column 0: DC or AC by the lane's row; other columns AC (uniform)
      const int32_t l = c == 0 ? quant_fpl<BD, Q24, false>(a, r != 0, qs.q[q], &e)
                               : quant_fpl<BD, Q24, true>(a, 1, qs.q[q], &e);
      lsum[q] += l;
      m[q] = max(m[q], l ? p : -1);
      sq4[q] = (uint32_t)mad_i24(e, e, (int32_t)sq4[q]);
    }
    if ((c & 3) == 3) {
#pragma unroll
      for (int q = 0; q < NQ; q++) dist[q] += sq4[q], sq4[q] = 0;
    }
  }
}

// Rows per lane of the row form, by measurement: 2 for 8x16 (its 8 lanes per
// block then each own a column in the column pass), 1 otherwise (two rows per
// lane measured no faster for the 16-wide rows, and costs LDS).
template <int W, int H>
constexpr int ts_rows_r() {
  return (W == 8 && H == 16) ? 2 : 1;
}

// Column pass of one variant: the lane's columns c = t + LP k (packed
// residual words, laundered per variant so the pass stays in the loop) into
// the block's LDS tile.
template <int W, int H, int CK, int LP, int NC>
__device__ __forceinline__ void ts_row_cols(const uint32_t (&rp)[NC][H / 2], int t, int32_t* __restrict__ tile) {
  constexpr int CBC = fwd_cb_col(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr long long kBCol = fwd_col_bound(SH.s0);
  constexpr int LS = W + 1;
#pragma unroll
  for (int k = 0; k < NC; k++) {
    int32_t wv[H / 2];
#pragma unroll
    for (int i = 0; i < H / 2; i++) wv[i] = (int32_t)rp[k][i];
    if constexpr (H / 2 >= 4) opaque(wv);
    else asm volatile("" : "+v"(wv[0]), "+v"(wv[1]));
    int32_t x[H];
#pragma unroll
    for (int rr = 0; rr < H; rr++) {
      const int s = CK == 2 ? H - 1 - rr : rr;  // FLIPADST: up-down flip
      x[rr] = (s & 1) ? (wv[s >> 1] >> 16) : (int32_t)(int16_t)(wv[s >> 1] & 0xffff);
    }
    fwd1d<H, CBC, kBCol>(x, CK);
    const int c = t + LP * k;
    if (c < W) {
#pragma unroll
      for (int rr = 0; rr < H; rr++) tile[rr * LS + c] = rshift_round32(x[rr], SH.s1);
    }
  }
}

template <int W, int H, int BD, bool Q24, int NQ>
__global__ __launch_bounds__(256) void k_txsearch_rows(const uint16_t* __restrict__ plane, int stride, int bd,
                                                       const int32_t* __restrict__ xy, int n_blocks, TxPlan plan,
                                                       QuantFpSet qs,
                                                       const uint16_t* __restrict__ iscan_all, int tx_size,
                                                       uint64_t* __restrict__ out_dist,
                                                       int32_t* __restrict__ out_eob,
                                                       uint32_t* __restrict__ out_lsum) {
  static_assert((W == 16 || H == 16) && W >= 8 && W * H <= 256, "row-form sizes");
  constexpr int R = ts_rows_r<W, H>();       // rows per lane
  constexpr int LP = H / R;                  // lanes per block
  constexpr int BPW = 64 / LP;               // blocks per wave
  constexpr int LS = W + 1;                  // padded tile row
  constexpr int NC = W >= LP ? W / LP : 1;   // columns per lane in the column pass
  constexpr FwdShift SH = fwd_shift(W, H);
  static_assert(4 * (BPW * H * LS * 4 + NQ * BPW * kStageRow * 16) <= 65536, "LDS of a 4-wave workgroup");
  __shared__ int32_t tiles[4][BPW * H * LS];
  __shared__ uint64_t st_d[NQ][4][BPW * kStageRow];
  __shared__ int32_t st_e[NQ][4][BPW * kStageRow];
  __shared__ uint32_t st_l[NQ][4][BPW * kStageRow];

  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = lane / LP, t = lane - g * LP;  // block within the wave, lane within the block
  const long bw0 = ((long)xcd_chunked(blockIdx.x, gridDim.x) * 4 + wv) * BPW;
  const long b = bw0 + g;
  const int bi = b < n_blocks ? (int)b : 0;  // past the end: block 0, never copied out
  const int x0 = xy[2 * bi], y0 = xy[2 * bi + 1];
  const uint32_t o0 = (uint32_t)y0 * (uint32_t)stride + (uint32_t)x0;
  const PlaneBuf pb(plane);
  const uint32_t rb = 2u * (uint32_t)stride;  // row pitch in bytes

  // DC prediction from the source neighbours (the block's LP lanes reduce)
  uint32_t sa = 0, sl = 0;
  if (y0 > 0) {
#pragma unroll
    for (int k = 0; k < NC; k++)
      if (t + LP * k < W) sa += (uint32_t)pb.at(2u * (o0 - (uint32_t)stride + (uint32_t)(t + LP * k)), 0);
  }
  if (x0 > 0) {
#pragma unroll
    for (int i = 0; i < R; i++) sl += (uint32_t)pb.at(2u * (o0 - 1u), rb * (uint32_t)(t * R + i));
  }
  sa = (uint32_t)dpp_sum_i32<LP>((int32_t)sa);
  sl = (uint32_t)dpp_sum_i32<LP>((int32_t)sl);
  int32_t dc;
  if (y0 > 0 && x0 > 0) dc = (int32_t)((sa + sl + (W + H) / 2) / (W + H));
  else if (y0 > 0) dc = (int32_t)((sa + W / 2) / W);
  else if (x0 > 0) dc = (int32_t)((sl + H / 2) / H);
  else dc = 1 << (BD - 1);

  // the lane's columns of the residual, packed: (sample - dc) << s0 in 16-bit pairs
  uint32_t rp[NC][H / 2];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    const int c = (t + LP * k) % W;  // lanes without a column (W < LP) read a valid one, never stored
#pragma unroll
    for (int i = 0; i < H / 2; i++) {
      const int32_t lo = (pb.at(2u * (o0 + (uint32_t)c), rb * (2 * i)) - dc) << SH.s0;
      const int32_t hi = (pb.at(2u * (o0 + (uint32_t)c), rb * (2 * i + 1)) - dc) << SH.s0;
      rp[k][i] = ((uint32_t)lo & 0xffffu) | ((uint32_t)hi << 16);
    }
  }
  // scan index of each coefficient of the lane's rows, three classes, bytes
  uint32_t pw[R][3][W / 4];
#pragma unroll
  for (int i = 0; i < R; i++)
#pragma unroll
    for (int cls = 0; cls < 3; cls++) {
      const uint16_t* is = iscan_all + (tx_size * 3 + cls) * 1024 + (t * R + i) * W;
#pragma unroll
      for (int k = 0; k < W / 4; k++)
        pw[i][cls][k] = (uint32_t)is[4 * k] | (uint32_t)is[4 * k + 1] << 8 | (uint32_t)is[4 * k + 2] << 16 |
                        (uint32_t)is[4 * k + 3] << 24;
    }

  const int ntyp = plan.n_types;
  int32_t* tile = tiles[wv] + g * H * LS;
  for (int v = 0; v < plan.n_var; v++) {
    const uint32_t vi = plan.lane_var[v];
    const int ck = (int)(vi & 0xff);
    wave_sync();  // the previous variant's rows have been read
    if (ck == 0) ts_row_cols<W, H, 0, LP, NC>(rp, t, tile);
    else if (ck == 1) ts_row_cols<W, H, 1, LP, NC>(rp, t, tile);
    else if (ck == 2) ts_row_cols<W, H, 2, LP, NC>(rp, t, tile);
    else ts_row_cols<W, H, 3, LP, NC>(rp, t, tile);
    wave_sync();
    const int i_end = (int)(vi >> 16);
    for (int i = (int)((vi >> 8) & 0xff); i < i_end; i++) {
      const uint32_t ti = plan.lane_type[i];
      const int cls = (int)((ti >> 16) & 0xff), rk = (int)((ti >> 8) & 0xff);
      uint64_t dist[NQ];
      int32_t m[NQ], lsum[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) dist[q] = 0, m[q] = -1, lsum[q] = 0;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int r = t * R + j;
        uint32_t pc[W / 4];
#pragma unroll
        for (int k = 0; k < W / 4; k++) pc[k] = cls == 0 ? pw[j][0][k] : (cls == 1 ? pw[j][1][k] : pw[j][2][k]);
        const int32_t* src = tile + r * LS;
        if (rk == 0) ts_row_arm<W, H, 0, BD, Q24, NQ>(src, pc, r, qs, dist, m, lsum);
        else if (rk == 1) ts_row_arm<W, H, 1, BD, Q24, NQ>(src, pc, r, qs, dist, m, lsum);
        else if (rk == 2) ts_row_arm<W, H, 2, BD, Q24, NQ>(src, pc, r, qs, dist, m, lsum);
        else ts_row_arm<W, H, 3, BD, Q24, NQ>(src, pc, r, qs, dist, m, lsum);
      }
      const int k = (int)(ti & 0xff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const uint64_t dq = dpp_sum_u64<LP>(dist[q]);
        const int32_t mq = dpp_max_i32<LP>(m[q]);
        const int32_t lq = dpp_sum_i32<LP>(lsum[q]);
        if (t == 0) {
          st_d[q][wv][g * kStageRow + k] = dq >> 2;  // (1 - log scale) * 2 with log scale 0
          st_e[q][wv][g * kStageRow + k] = mq + 1;
          st_l[q][wv][g * kStageRow + k] = (uint32_t)lq;
        }
      }
    }
  }
  wave_sync();

  // the wave's blocks own one contiguous range of each output array
  const long nb = n_blocks - bw0 < BPW ? n_blocks - bw0 : BPW;
  const uint32_t mg = (65536u + (uint32_t)ntyp - 1u) / (uint32_t)ntyp;
  for (int j = lane; j < nb * ntyp; j += 64) {
    const int blk = (int)(((uint32_t)j * mg) >> 16), k = j - blk * ntyp;
    const int src = blk * kStageRow + k;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const long o = (long)q * n_blocks * ntyp + bw0 * ntyp + j;
      out_dist[o] = st_d[q][wv][src];
      out_eob[o] = st_e[q][wv][src];
      out_lsum[o] = st_l[q][wv][src];
    }
  }
  (void)bd;
}

// Sizes evaluated by the row form.
constexpr bool ts_row_size(int w, int h) { return (w == 16 || h == 16) && w >= 8 && w * h <= 256; }

// Sizes evaluated by the lane-per-block kernel (by measurement: 16-wide rows
// hold too many registers per lane; 16x4 stays on the group form).
constexpr bool ts_lane_size(int w, int h) { return w * h <= 64 && w <= 8; }

// Lanes per block, picked per size by measurement on MI355X (config-2
// sweep): more lanes than max(W, H) only where the extra type slots of the
// row pass beat the per-block overhead they add.
constexpr int ts_lanes(int w, int h) {
  return (w == 4 && h == 4) ? 8
       : (w == 8 && h == 8) ? 32
       : (w == 16 && h == 16) ? 64
       : ((w == 8 && h == 16) || (w == 16 && h == 8)) ? 32
       : (w > h ? w : h);
}
static TxPlan plan_by_variant(const TxPlan& a);

// Fused (lane / row form) launch for quantizers [q0, q0 + nq) of the list,
// nq <= 2, at the bit depth and level form of the rows.
template <int W, int H, template <int, int, int, bool, int> class K, typename... Args>
static void launch_fused(int bd, bool q24, int nq, dim3 grid, hipStream_t st, Args... args) {
  if (bd == 8) {
    if (nq == 2) hipLaunchKernelGGL((K<W, H, 8, true, 2>::fn), grid, dim3(256), 0, st, args...);
    else hipLaunchKernelGGL((K<W, H, 8, true, 1>::fn), grid, dim3(256), 0, st, args...);
  } else if (q24) {
    if (nq == 2) hipLaunchKernelGGL((K<W, H, 10, true, 2>::fn), grid, dim3(256), 0, st, args...);
    else hipLaunchKernelGGL((K<W, H, 10, true, 1>::fn), grid, dim3(256), 0, st, args...);
  } else {
    if (nq == 2) hipLaunchKernelGGL((K<W, H, 10, false, 2>::fn), grid, dim3(256), 0, st, args...);
    else hipLaunchKernelGGL((K<W, H, 10, false, 1>::fn), grid, dim3(256), 0, st, args...);
  }
}
template <int W, int H, int BD, bool Q24, int NQ>
struct LaneK {
  static constexpr auto fn = k_txsearch_lane<W, H, BD, Q24, NQ>;
};
template <int W, int H, int BD, bool Q24, int NQ>
struct RowK {
  static constexpr auto fn = k_txsearch_rows<W, H, BD, Q24, NQ>;
};

template <int S>
static hipError_t launch_ts(const uint16_t* plane, int stride, int bd, const int32_t* xy, int n_blocks,
                            const TxPlan& plan, const QuantRow* const* dev_rows, const QuantRow* host_rows, int n_q,
                            const uint16_t* iscan, uint64_t* d, int32_t* e, uint32_t* l, hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const long per_q = (long)n_blocks * plan.n_types;
  if constexpr (ts_lane_size(W, H) || ts_row_size(W, H)) {
    const TxPlan lp = plan_by_variant(plan);
    LaneScan sc{};
    if constexpr (ts_lane_size(W, H)) {
      uint16_t order[1024];
      for (int c = 0; c < 3; c++) {
        build_scan(S, c, order);
        for (int i = 0; i < W * H; i++) sc.pos[c][order[i] >> 2] |= (uint32_t)(i + 1) << (8 * (order[i] & 3));
      }
    }
    for (int q0 = 0; q0 < n_q; q0 += 2) {
      const int nq = n_q - q0 < 2 ? n_q - q0 : 2;
      bool q24 = true;  // the 24-bit level form (exact when quant_fp < 2^14 at 10 bits)
      QuantFpSet qs{};
      for (int i = 0; i < nq; i++)
        q24 = q24 && (bd == 8 || (host_rows[q0 + i].quant_fp[0] < 16384 && host_rows[q0 + i].quant_fp[1] < 16384));
      for (int i = 0; i < nq; i++) qs.q[i] = make_quant_fpl(host_rows[q0 + i], bd, q24);
      uint64_t* dq = d + q0 * per_q;
      int32_t* eq = e + q0 * per_q;
      uint32_t* lq = l + q0 * per_q;
      if constexpr (ts_lane_size(W, H)) {
        // one wave per column variant: 4 waves (the plans here have <= 4 variants)
        const dim3 grid((n_blocks + kLaneBlocks - 1) / kLaneBlocks);
        launch_fused<W, H, LaneK>(bd, q24, nq, grid, st, plane, stride, bd, xy, n_blocks, lp, qs, sc, dq, eq, lq);
      } else {
        constexpr int BPW = 64 / (H / ts_rows_r<W, H>());  // blocks per wave
        const dim3 grid((n_blocks + 4 * BPW - 1) / (4 * BPW));
        launch_fused<W, H, RowK>(bd, q24, nq, grid, st, plane, stride, bd, xy, n_blocks, lp, qs, iscan, S, dq, eq,
                                 lq);
      }
    }
  } else {
    constexpr int P = ts_lanes(W, H);
    using G = BlkGeom<W, H, P>;
    // groups per workgroup: up to 256 threads and 64 KB of column variants
    const int per_group = plan.n_var * G::VS * (int)sizeof(int32_t);
    int groups = 256 / G::T;
    while (groups > 1 && groups * per_group > 65536) groups >>= 1;
    const int grid = (n_blocks + groups - 1) / groups;
    constexpr int LSC = tx_log_scale(W, H);
    for (int q0 = 0; q0 < n_q; q0 += 2) {
      const int nq = n_q - q0 < 2 ? n_q - q0 : 2;
      QuantFpPair qp{};
      for (int i = 0; i < nq; i++) qp.q[i] = make_quant_fp(host_rows[q0 + i], LSC);
      uint64_t* dq = d + q0 * per_q;
      int32_t* eq = e + q0 * per_q;
      uint32_t* lq = l + q0 * per_q;
      const dim3 gr(grid), bl(groups * G::T);
      const size_t sh = (size_t)groups * per_group;
      if (bd == 8) {
        if (nq == 2)
          hipLaunchKernelGGL((k_txsearch<W, H, P, 8, 2>), gr, bl, sh, st, plane, stride, bd, xy, n_blocks, plan, qp,
                             iscan, S, dq, eq, lq);
        else
          hipLaunchKernelGGL((k_txsearch<W, H, P, 8, 1>), gr, bl, sh, st, plane, stride, bd, xy, n_blocks, plan, qp,
                             iscan, S, dq, eq, lq);
      } else {
        if (nq == 2)
          hipLaunchKernelGGL((k_txsearch<W, H, P, 10, 2>), gr, bl, sh, st, plane, stride, bd, xy, n_blocks, plan, qp,
                             iscan, S, dq, eq, lq);
        else
          hipLaunchKernelGGL((k_txsearch<W, H, P, 10, 1>), gr, bl, sh, st, plane, stride, bd, xy, n_blocks, plan, qp,
                             iscan, S, dq, eq, lq);
      }
    }
    (void)dev_rows;
  }
  return hipGetLastError();
}

// AV1 coefficient scans of the packed (<= 32x32) block: class 0 runs the
// anti-diagonals (square blocks alternate direction starting up-right, wide
// blocks always run up-right, tall blocks down-left), class 1 is raster,
// class 2 column-major.  Same orders as the C-extracted tables of
// rust/crates/svtav1-entropy/src/scan_tables.rs.
void build_scan(int tx_size, int cls, uint16_t* scan) {
  int w = tx_w(tx_size), h = tx_h(tx_size);
  w = w > 32 ? 32 : w;
  h = h > 32 ? 32 : h;
  int k = 0;
  if (cls == 1) {
    for (int i = 0; i < w * h; i++) scan[k++] = (uint16_t)i;
  } else if (cls == 2) {
    for (int c = 0; c < w; c++)
      for (int r = 0; r < h; r++) scan[k++] = (uint16_t)(r * w + c);
  } else {
    for (int d = 0; d < w + h - 1; d++) {
      const bool up = (w == h) ? ((d & 1) == 0) : (w > h);
      for (int i = 0; i < h; i++) {
        const int r = up ? (h - 1 - i) : i, c = d - r;
        if (c >= 0 && c < w) scan[k++] = (uint16_t)(r * w + c);
      }
    }
  }
}

// Processing plan of a type list: types grouped by row kind (ADST and
// FLIPADST rows share a kernel), distinct column variants (kind, up-down
// flip) numbered in first-use order.
static TxPlan make_plan(const TypeList& types) {
  TxPlan p{};
  p.n_types = types.n;
  int order[16], n = 0;
  for (int key = 0; key < 3; key++)
    for (int i = 0; i < types.n; i++) {
      const int rk = type_cfg(types.t[i]).row;
      const int k2 = rk == 0 ? 0 : (rk == 3 ? 2 : 1);
      if (k2 == key) order[n++] = i;
    }
  for (int i = 0; i < n; i++) {
    const int k = order[i];
    const TypeCfg c = type_cfg(types.t[k]);
    const int kind = c.col == 2 ? 1 : c.col;
    int v = 0;
    while (v < p.n_var && !(p.var_kind[v] == kind && p.var_ud[v] == c.ud)) v++;
    if (v == p.n_var) {
      p.var_kind[v] = (uint8_t)kind;
      p.var_ud[v] = (uint8_t)c.ud;
      p.n_var++;
    }
    p.type[i] = types.t[k];
    p.out_k[i] = (uint8_t)k;
    p.var_of[i] = (uint8_t)v;
    p.row_kind[i] = (uint8_t)c.row;
    const int t = types.t[k];
    p.scan_cls[i] = (uint8_t)((t >= 10) ? (((t - 10) & 1) ? 2 : 1) : 0);
  }
  return p;
}

// The same plan with its types grouped by column variant (stable), for the
// lane-per-block kernel.
static TxPlan plan_by_variant(const TxPlan& a) {
  TxPlan p = a;
  int n = 0;
  for (int v = 0; v < a.n_var; v++) {
    p.var_start[v] = (uint8_t)n;
    for (int i = 0; i < a.n_types; i++)
      if (a.var_of[i] == v) {
        p.type[n] = a.type[i];
        p.out_k[n] = a.out_k[i];
        p.var_of[n] = a.var_of[i];
        p.row_kind[n] = a.row_kind[i];
        p.scan_cls[n] = a.scan_cls[i];
        n++;
      }
  }
  p.var_start[a.n_var] = (uint8_t)n;
  for (int i = 0; i < n; i++)
    p.lane_type[i] = (uint32_t)p.out_k[i] | (uint32_t)p.row_kind[i] << 8 | (uint32_t)p.scan_cls[i] << 16;
  for (int v = 0; v < a.n_var; v++) {
    const uint32_t ck = p.var_ud[v] ? 2u : p.var_kind[v];
    p.lane_var[v] = ck | (uint32_t)p.var_start[v] << 8 | (uint32_t)p.var_start[v + 1] << 16;
  }
  return p;
}

hipError_t txsearch_batch(const uint16_t* plane, int stride, int bd, int tx_size, const int32_t* xy,
                          int n_blocks, const TypeList& types, const QuantRow* const* dev_rows,
                          const QuantRow* host_rows, int n_q, const uint16_t* iscan, uint64_t* dist, int32_t* eob,
                          uint32_t* lsum, hipStream_t st) {
  if (n_blocks <= 0 || types.n <= 0) return hipSuccess;
  const TxPlan plan = make_plan(types);
  switch (tx_size) {
#define X(S) \
  case S: return launch_ts<S>(plane, stride, bd, xy, n_blocks, plan, dev_rows, host_rows, n_q, iscan, dist, eob, lsum, st);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)
#undef X
    default: return hipErrorInvalidValue;
  }
}

__global__ void k_grid_xy(int32_t* __restrict__ xy, int nx, int n, int bw, int bh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int by = i / nx, bx = i - by * nx;
  xy[2 * i] = bx * bw;
  xy[2 * i + 1] = by * bh;
}

hipError_t grid_xy(int32_t* xy, int nx, int n, int bw, int bh, hipStream_t st) {
  hipLaunchKernelGGL(k_grid_xy, dim3((n + 255) / 256), dim3(256), 0, st, xy, nx, n, bw, bh);
  return hipGetLastError();
}

}  // namespace zav1
