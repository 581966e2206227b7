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

template <int W, int H>
struct TsGeom {
  static constexpr int T = W > H ? W : H;
  static constexpr int JPW = 256 / T;
  static constexpr int LS = W + 1;
  static constexpr int BLK = H * LS;
  static constexpr int PW = W < 32 ? W : 32, PH = H < 32 ? H : 32;
  static constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
};

// log_scale per TxSize (rust/crates/svtav1-encoder/src/quant.rs:238).
__host__ __device__ constexpr int tx_log_scale(int w, int h) {
  return (w * h > 1024) ? 2 : ((w * h > 256) ? 1 : 0);
}

template <int T>
__device__ __forceinline__ uint64_t group_sum_u64(uint64_t v) {
#pragma unroll
  for (int m = 1; m < T; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
template <int T>
__device__ __forceinline__ int32_t group_sum_i32(int32_t v) {
#pragma unroll
  for (int m = 1; m < T; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
template <int T>
__device__ __forceinline__ int32_t group_max_i32(int32_t v) {
#pragma unroll
  for (int m = 1; m < T; m <<= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}

template <int W, int H>
__global__ __launch_bounds__(256) void k_txsearch(const uint16_t* __restrict__ plane, int stride, int bd,
                                                  const int32_t* __restrict__ xy, int n_blocks,
                                                  TypeList types, const QuantRow* __restrict__ row_p,
                                                  const uint16_t* __restrict__ iscan_all, int tx_size,
                                                  uint64_t* __restrict__ out_dist,
                                                  int32_t* __restrict__ out_eob,
                                                  uint32_t* __restrict__ out_lsum) {
  using G = TsGeom<W, H>;
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  constexpr int LSC = tx_log_scale(W, H);
  __shared__ int32_t lds[G::JPW * G::BLK];

  const int jl = threadIdx.x / G::T, t = threadIdx.x - jl * G::T;
  const long pair = (long)blockIdx.x * G::JPW + jl;
  const long n_pairs = (long)n_blocks * types.n;
  const bool active = pair < n_pairs;
  const int blk_i = active ? (int)(pair / types.n) : 0;
  const int k = active ? (int)(pair - (long)blk_i * types.n) : 0;
  const int tx_type = types.t[k];
  const TypeCfg cfg = type_cfg(tx_type);
  const int x0 = xy[2 * blk_i], y0 = xy[2 * blk_i + 1];

  // ---- DC prediction from the source neighbours (group reduction) ----
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

  int32_t* blk = lds + jl * G::BLK;
  int32_t v[G::T];
  const bool do_col = active && t < W;
  if (do_col) {
#pragma unroll
    for (int r = 0; r < H; r++) {
      const int rr = cfg.ud ? (H - 1 - r) : r;
      v[r] = ((int32_t)plane[(long)(y0 + rr) * stride + x0 + t] - dc) << SH.s0;
    }
    fwd1d<H, CBC>(v, cfg.col);
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = rshift_round(v[r], SH.s1);
    const int cc = cfg.lr ? (W - 1 - t) : t;
#pragma unroll
    for (int r = 0; r < H; r++) blk[r * G::LS + cc] = v[r];
  }
  __syncthreads();

  uint64_t dist = 0;
  int32_t eob = 0, lsum = 0;
  if (active && t < H) {
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = blk[t * G::LS + c];
    fwd1d<W, CBR>(v, cfg.row);
    const QuantRow& qr = *row_p;
    const int cls = (tx_type >= 10) ? (((tx_type - 10) & 1) ? 2 : 1) : 0;
    const uint16_t* iscan = iscan_all + (tx_size * 3 + cls) * 1024;
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = rshift_round(v[c], SH.s2);
      if constexpr (G::RECT2) x = rshift64r((int64_t)x * 5793, 12);
      if (t >= 32 || c >= 32) {
        dist += (uint64_t)((int64_t)x * x);  // 64-point fold: discarded energy
        continue;
      }
      const int iz = (t | c) != 0;
      const int32_t sgn = x >> 31;
      const int32_t a = (x ^ sgn) - sgn;
      int32_t q = 0;
      const int32_t rounding = (qr.round_fp[iz] + ((1 << LSC) >> 1)) >> LSC;
      if (((int64_t)a << (1 + LSC)) >= qr.dequant[iz]) {
        int64_t xx = (int64_t)a + rounding;
        if (bd == 8) xx = xx > 32767 ? 32767 : xx;
        q = (int32_t)((xx * qr.quant_fp[iz]) >> (16 - LSC));
      }
      int32_t dq = 0;
      if (q != 0) {
        dq = (int32_t)(((int64_t)q * qr.dequant[iz]) >> LSC);
        lsum += q;
        eob = max(eob, (int32_t)iscan[t * G::PW + c] + 1);
        dq = (dq ^ sgn) - sgn;
      }
      const int64_t e = (int64_t)x - dq;
      dist += (uint64_t)e * (uint64_t)e;
    }
  }
  dist = group_sum_u64<G::T>(dist);
  eob = group_max_i32<G::T>(eob);
  lsum = group_sum_i32<G::T>(lsum);
  if (active && t == 0) {
    constexpr int SHF = (1 - LSC) * 2;
    out_dist[pair] = SHF < 0 ? (dist << (-SHF)) : (dist >> SHF);
    out_eob[pair] = eob;
    out_lsum[pair] = (uint32_t)lsum;
  }
}

template <int S>
static hipError_t launch_ts(const uint16_t* plane, int stride, int bd, const int32_t* xy, int n_blocks,
                            const TypeList& types, const QuantRow* row, const uint16_t* iscan,
                            uint64_t* d, int32_t* e, uint32_t* l, hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const long pairs = (long)n_blocks * types.n;
  const int grid = (int)((pairs + TsGeom<W, H>::JPW - 1) / TsGeom<W, H>::JPW);
  hipLaunchKernelGGL((k_txsearch<W, H>), dim3(grid), dim3(256), 0, st, plane, stride, bd, xy, n_blocks,
                     types, row, iscan, S, d, e, l);
  return hipGetLastError();
}

hipError_t txsearch_batch(const uint16_t* plane, int stride, int bd, int tx_size, const int32_t* xy,
                          int n_blocks, const TypeList& types, const QuantRow* row, const uint16_t* iscan,
                          uint64_t* dist, int32_t* eob, uint32_t* lsum, hipStream_t st) {
  if (n_blocks <= 0 || types.n <= 0) return hipSuccess;
  switch (tx_size) {
#define X(S) case S: return launch_ts<S>(plane, stride, bd, xy, n_blocks, types, row, iscan, dist, eob, lsum, st);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)
#undef X
    default: return hipErrorInvalidValue;
  }
}

}  // namespace zav1
