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

template <int W, int H, int P, int BD>
__global__ __launch_bounds__(256) void k_txsearch(const uint16_t* __restrict__ plane, int stride, int bd,
                                                  const int32_t* __restrict__ xy, int n_blocks, TxPlan plan,
                                                  const QuantRow* __restrict__ row_p,
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

  // ---- row passes: items (type, row) ----
  constexpr int TPR = G::T / H;
  const int sub = t / H, r = t - sub * H;
  const QuantRow& qr = *row_p;
  const QuantFp qf = make_quant_fp(qr, LSC);
  for (int base = 0; base < ntyp; base += TPR) {
    const int ti = base + sub;
    const bool live = active && ti < ntyp;
    uint64_t dist = 0;
    int32_t eob = 0, lsum = 0;
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
      uint32_t sq4 = 0;
#pragma unroll
      for (int c = 0; c < W; c++) {
        int32_t x = rshift_round32(v[c], SH.s2);
        // 2:1 rectangles: x NewSqrt2 / 2^12; |x| stays below 2^18 (the pinned
        // coefficient bound before this scaling), so the product fits 32 bits
        if constexpr (G::RECT2) x = (__mul24(x, 5793) + 2048) >> 12;
        if (r >= 32 || c >= 32) {
          dist += (uint64_t)((int64_t)x * x);  // 64-point fold: discarded energy
          continue;
        }
        const int32_t a = x < 0 ? -x : x;
        int32_t adq;
        const int32_t q = quant_fp_fast<LSC>(a, (r | c) != 0, qf, BD, &adq);
        lsum += q;
        eob = max(eob, q ? sp[c < KW ? c : 0] : 0);
        // x - dq = sign(x) * (a - |dq|): the squared error needs magnitudes only
        const int32_t e = a - adq;
        sq4 += (uint32_t)__mul24(e, e);
        if ((c & 3) == 3 || c == W - 1 || c == 31) {
          dist += sq4;
          sq4 = 0;
        }
      }
    }
    dist = group_sum_u64<H>(dist);
    eob = group_max_i32<H>(eob);
    lsum = group_sum_i32<H>(lsum);
    if (live && r == 0) {
      constexpr int SHF = (1 - LSC) * 2;
      const long o = (long)bi * ntyp + plan.out_k[ti];
      out_dist[o] = SHF < 0 ? (dist << (-SHF)) : (dist >> SHF);
      out_eob[o] = eob;
      out_lsum[o] = (uint32_t)lsum;
    }
  }
}

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
template <int S>
static hipError_t launch_ts(const uint16_t* plane, int stride, int bd, const int32_t* xy, int n_blocks,
                            const TypeList& types, const TxPlan& plan, const QuantRow* row, const uint16_t* iscan,
                            uint64_t* d, int32_t* e, uint32_t* l, hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  constexpr int P = ts_lanes(W, H);
  using G = BlkGeom<W, H, P>;
  // groups per workgroup: up to 256 threads and 64 KB of column variants
  const int per_group = plan.n_var * G::VS * (int)sizeof(int32_t);
  int groups = 256 / G::T;
  while (groups > 1 && groups * per_group > 65536) groups >>= 1;
  const int grid = (n_blocks + groups - 1) / groups;
  if (bd == 8)
    hipLaunchKernelGGL((k_txsearch<W, H, P, 8>), dim3(grid), dim3(groups * G::T), groups * per_group, st, plane,
                       stride, bd, xy, n_blocks, plan, row, iscan, S, d, e, l);
  else
    hipLaunchKernelGGL((k_txsearch<W, H, P, 10>), dim3(grid), dim3(groups * G::T), groups * per_group, st, plane,
                       stride, bd, xy, n_blocks, plan, row, iscan, S, d, e, l);
  (void)types;
  return hipGetLastError();
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
  }
  return p;
}

hipError_t txsearch_batch(const uint16_t* plane, int stride, int bd, int tx_size, const int32_t* xy,
                          int n_blocks, const TypeList& types, const QuantRow* row, const uint16_t* iscan,
                          uint64_t* dist, int32_t* eob, uint32_t* lsum, hipStream_t st) {
  if (n_blocks <= 0 || types.n <= 0) return hipSuccess;
  const TxPlan plan = make_plan(types);
  switch (tx_size) {
#define X(S) case S: return launch_ts<S>(plane, stride, bd, xy, n_blocks, types, plan, row, iscan, dist, eob, lsum, st);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)
#undef X
    default: return hipErrorInvalidValue;
  }
}

}  // namespace zav1
