// Batched 2D forward / inverse AV1 transforms for gfx950.
//
// One launch handles a batch of blocks of ONE transform size (each block may
// carry its own tx type).  A 256-thread workgroup owns JPW = 256 / max(w,h)
// blocks; every block gets max(w,h) lanes.  Blocks are staged through LDS
// with a padded row stride (w + 1) so both the column walk (lane = column)
// and the row walk (lane = row) are bank-conflict free; global loads and
// stores are whole-workgroup coalesced copies of the packed blocks.
//
// Semantics (rounding points, shifts, flips, rect scaling, 64-point inverse
// zero-extension and bd clamps) follow
//   rust/crates/svtav1-dsp/src/fwd_txfm.rs:1891 fwd_txfm2d_core
//   rust/crates/svtav1-dsp/src/inv_txfm.rs:1816 inv_txfm2d_core
//   rust/crates/svtav1-dsp/src/txfm_dispatch.rs:18,55
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

template <int W, int H>
struct TxGeom {
  static constexpr int T = W > H ? W : H;        // lanes per block
  static constexpr int JPW = 256 / T;            // blocks per workgroup
  static constexpr int LS = W + 1;               // padded LDS row stride
  static constexpr int BLK = H * LS;             // LDS words per block
  static constexpr int N = W * H;
  static constexpr bool RECT2 = (W == 2 * H) || (H == 2 * W);
};

template <int W, int H>
__global__ __launch_bounds__(256) void k_fwd_txfm2d(const int32_t* __restrict__ in,
                                                    int32_t* __restrict__ out,
                                                    const uint8_t* __restrict__ types, int n_jobs) {
  using G = TxGeom<W, H>;
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift SH = fwd_shift(W, H);
  __shared__ int32_t lds[G::JPW * G::BLK];

  const int job0 = blockIdx.x * G::JPW;
  const int njobs_here = min(G::JPW, n_jobs - job0);
  const int32_t* src = in + (size_t)job0 * G::N;
  for (int i = threadIdx.x; i < njobs_here * G::N; i += 256) {
    const int j = i / G::N, e = i - j * G::N;
    lds[j * G::BLK + (e / W) * G::LS + (e % W)] = src[i];
  }
  __syncthreads();

  const int jl = threadIdx.x / G::T, t = threadIdx.x - jl * G::T;
  const bool active = jl < njobs_here;
  const TypeCfg cfg = type_cfg(active ? types[job0 + jl] : 0);
  int32_t* blk = lds + jl * G::BLK;
  int32_t v[G::T];

  // column pass: lane t < W transforms column t
  const bool do_col = active && t < W;
  if (do_col) {
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = blk[(cfg.ud ? (H - 1 - r) : r) * G::LS + t] << SH.s0;
    fwd1d<H, CBC>(v, cfg.col);
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = rshift_round(v[r], SH.s1);
  }
  __syncthreads();
  if (do_col) {
    const int cc = cfg.lr ? (W - 1 - t) : t;
#pragma unroll
    for (int r = 0; r < H; r++) blk[r * G::LS + cc] = v[r];
  }
  __syncthreads();

  // row pass: lane t < H transforms row t
  const bool do_row = active && t < H;
  if (do_row) {
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = blk[t * G::LS + c];
    fwd1d<W, CBR>(v, cfg.row);
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = rshift_round(v[c], SH.s2);
      if constexpr (G::RECT2) x = rshift64r((int64_t)x * 5793, 12);
      v[c] = x;
    }
  }
  __syncthreads();
  if (do_row) {
#pragma unroll
    for (int c = 0; c < W; c++) blk[t * G::LS + c] = v[c];
  }
  __syncthreads();

  int32_t* dst = out + (size_t)job0 * G::N;
  for (int i = threadIdx.x; i < njobs_here * G::N; i += 256) {
    const int j = i / G::N, e = i - j * G::N;
    dst[i] = lds[j * G::BLK + (e / W) * G::LS + (e % W)];
  }
}

// Inverse: coefficients (w x h raster, only the top-left 32x32 read for
// 64-point sizes) -> residual, HIGHBD_WRAPLOW clamp at bd.
template <int W, int H>
__global__ __launch_bounds__(256) void k_inv_txfm2d(const int32_t* __restrict__ in,
                                                    int32_t* __restrict__ out,
                                                    const uint8_t* __restrict__ types, int n_jobs,
                                                    int bd) {
  using G = TxGeom<W, H>;
  constexpr InvShift SH = inv_shift(W, H);
  __shared__ int32_t lds[G::JPW * G::BLK];

  const int job0 = blockIdx.x * G::JPW;
  const int njobs_here = min(G::JPW, n_jobs - job0);
  const int32_t* src = in + (size_t)job0 * G::N;
  for (int i = threadIdx.x; i < njobs_here * G::N; i += 256) {
    const int j = i / G::N, e = i - j * G::N;
    const int r = e / W, c = e % W;
    lds[j * G::BLK + r * G::LS + c] = (r < 32 && c < 32) ? src[i] : 0;
  }
  __syncthreads();

  const int jl = threadIdx.x / G::T, t = threadIdx.x - jl * G::T;
  const bool active = jl < njobs_here;
  const TypeCfg cfg = type_cfg(active ? types[job0 + jl] : 0);
  int32_t* blk = lds + jl * G::BLK;
  int32_t v[G::T];
  const int row_rng = bd + 8, col_rng = (bd + 6) > 16 ? bd + 6 : 16;
  const int32_t wrap_max = (1 << (7 + bd)) - 1 + (914 << (bd - 7));

  const bool do_row = active && t < H;
  if (do_row) {
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t x = blk[t * G::LS + c];
      if constexpr (G::RECT2) x = rshift64r((int64_t)x * 2896, 12);
      v[c] = zclamp(x, row_rng);
    }
    inv1d<W>(v, cfg.row, row_rng);
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = rshift_round(v[c], SH.r0);
  }
  __syncthreads();
  if (do_row) {
#pragma unroll
    for (int c = 0; c < W; c++) blk[t * G::LS + c] = v[c];
  }
  __syncthreads();

  const bool do_col = active && t < W;
  if (do_col) {
    const int cc = cfg.lr ? (W - 1 - t) : t;
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = zclamp(blk[r * G::LS + cc], col_rng);
    inv1d<H>(v, cfg.col, col_rng);
#pragma unroll
    for (int r = 0; r < H; r++) {
      int32_t x = rshift_round(v[r], SH.r1);
      v[r] = x < -wrap_max - 1 ? -wrap_max - 1 : (x > wrap_max ? wrap_max : x);
    }
  }
  __syncthreads();
  if (do_col) {
#pragma unroll
    for (int r = 0; r < H; r++) blk[(cfg.ud ? (H - 1 - r) : r) * G::LS + t] = v[r];
  }
  __syncthreads();

  int32_t* dst = out + (size_t)job0 * G::N;
  for (int i = threadIdx.x; i < njobs_here * G::N; i += 256) {
    const int j = i / G::N, e = i - j * G::N;
    dst[i] = lds[j * G::BLK + (e / W) * G::LS + (e % W)];
  }
}

template <int S>
static hipError_t launch_fwd(const int32_t* in, int32_t* out, const uint8_t* types, int n, hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const int grid = (n + TxGeom<W, H>::JPW - 1) / TxGeom<W, H>::JPW;
  hipLaunchKernelGGL((k_fwd_txfm2d<W, H>), dim3(grid), dim3(256), 0, st, in, out, types, n);
  return hipGetLastError();
}
template <int S>
static hipError_t launch_inv(const int32_t* in, int32_t* out, const uint8_t* types, int n, int bd,
                             hipStream_t st) {
  constexpr int W = tx_w(S), H = tx_h(S);
  const int grid = (n + TxGeom<W, H>::JPW - 1) / TxGeom<W, H>::JPW;
  hipLaunchKernelGGL((k_inv_txfm2d<W, H>), dim3(grid), dim3(256), 0, st, in, out, types, n, bd);
  return hipGetLastError();
}

#define ZAV1_SIZE_SWITCH(X)                                                                     \
  switch (tx_size) {                                                                            \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)      \
    X(16) X(17) X(18)                                                                           \
    default: return hipErrorInvalidValue;                                                       \
  }

hipError_t fwd_txfm2d_batch(const int32_t* in, int32_t* out, const uint8_t* types, int tx_size,
                            int n_jobs, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
#define X(S) case S: return launch_fwd<S>(in, out, types, n_jobs, st);
  ZAV1_SIZE_SWITCH(X)
#undef X
}

hipError_t inv_txfm2d_batch(const int32_t* in, int32_t* out, const uint8_t* types, int tx_size,
                            int n_jobs, int bd, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
#define X(S) case S: return launch_inv<S>(in, out, types, n_jobs, bd, st);
  ZAV1_SIZE_SWITCH(X)
#undef X
}

}  // namespace zav1
