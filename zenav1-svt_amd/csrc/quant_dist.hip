// Batched quantizers, frequency-domain SSE and Hadamard SATD for gfx950.
//
//   quantize_fp / quantize_b   rust/crates/svtav1-dsp/src/quant_coding.rs:80,365
//                              (bd8 INT16 clamp; bd10 per
//                              rust/crates/svtav1-encoder/src/quant.rs:335,391)
//   sse_i32 / sq_sum_i32       rust/crates/svtav1-dsp/src/residual.rs:245,364
//   hadamard 4/8/16/32 + satd  rust/crates/svtav1-dsp/src/hadamard.rs:643-948
#include "hadamard.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

// ---------------------------------------------------------------------------
// quantizers: one lane per coefficient, rows from a small device table.
// ---------------------------------------------------------------------------
template <int KIND, int BD>
__global__ __launch_bounds__(256) void k_quantize(const int32_t* __restrict__ coeff,
                                                  int32_t* __restrict__ qcoeff,
                                                  int32_t* __restrict__ dqcoeff,
                                                  const QuantRow* __restrict__ rows,
                                                  const uint8_t* __restrict__ qidx, int n_per_job,
                                                  long total, int ls) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const int job = (int)(i / n_per_job);
  const int rc = (int)(i - (long)job * n_per_job);
  const int iz = rc != 0;
  const QuantRow& t = rows[qidx[job]];
  const int32_t c = coeff[i];
  const int32_t sgn = c >> 31;
  const int32_t a = (c ^ sgn) - sgn;
  const int32_t q = KIND == 0 ? quant_fp_level(a, t, iz, ls, BD) : quant_b_level(a, t, iz, ls, BD);
  int32_t qo = 0, dqo = 0;
  if (q != 0) {
    qo = (q ^ sgn) - sgn;
    const int32_t adq = dequant_level(q, t, iz, ls);
    dqo = (adq ^ sgn) - sgn;
  }
  qcoeff[i] = qo;
  dqcoeff[i] = dqo;
}

hipError_t quantize_batch(const int32_t* coeff, int32_t* qcoeff, int32_t* dqcoeff,
                          const QuantRow* rows, const uint8_t* qidx, int n_per_job, int n_jobs,
                          int log_scale, int bd, int kind, hipStream_t st) {
  const long total = (long)n_per_job * n_jobs;
  if (total <= 0) return hipSuccess;
  const int grid = (int)((total + 255) / 256);
  if (kind == 0 && bd == 8)
    hipLaunchKernelGGL((k_quantize<0, 8>), dim3(grid), dim3(256), 0, st, coeff, qcoeff, dqcoeff, rows, qidx, n_per_job, total, log_scale);
  else if (kind == 0)
    hipLaunchKernelGGL((k_quantize<0, 10>), dim3(grid), dim3(256), 0, st, coeff, qcoeff, dqcoeff, rows, qidx, n_per_job, total, log_scale);
  else if (bd == 8)
    hipLaunchKernelGGL((k_quantize<1, 8>), dim3(grid), dim3(256), 0, st, coeff, qcoeff, dqcoeff, rows, qidx, n_per_job, total, log_scale);
  else
    hipLaunchKernelGGL((k_quantize<1, 10>), dim3(grid), dim3(256), 0, st, coeff, qcoeff, dqcoeff, rows, qidx, n_per_job, total, log_scale);
  return hipGetLastError();
}

// Plain-form quantizer: caller-supplied, already log_scale-adjusted rows
// (the arguments of quantize_fp_raster / quantize_b_raster), bd8 clamp.
__global__ __launch_bounds__(256) void k_quantize_plain(const int32_t* __restrict__ coeff,
                                                        int32_t* __restrict__ qcoeff,
                                                        int32_t* __restrict__ dqcoeff, int n,
                                                        PlainQuant p) {
  const int rc = blockIdx.x * 256 + threadIdx.x;
  if (rc >= n) return;
  const int iz = rc != 0;
  const int32_t c = coeff[rc];
  const int32_t sgn = c >> 31;
  const int32_t a = (c ^ sgn) - sgn;
  int32_t q = 0;
  if (p.kind == 0) {
    if (((int64_t)a << (1 + p.log_scale)) >= p.dequant[iz]) {
      int64_t x = (int64_t)a + p.round[iz];
      x = x > 32767 ? 32767 : x;
      q = (int32_t)((x * p.quant[iz]) >> (16 - p.log_scale));
    }
  } else {
    if (a >= p.zbin[iz]) {
      int64_t x = (int64_t)a + p.round[iz];
      x = x > 32767 ? 32767 : x;
      q = (int32_t)(((((x * p.quant[iz]) >> 16) + x) * p.quant_shift[iz]) >> (16 - p.log_scale));
    }
  }
  int32_t qo = 0, dqo = 0;
  if (q != 0) {
    qo = (q ^ sgn) - sgn;
    const int32_t adq = (int32_t)(((int64_t)q * p.dequant[iz]) >> p.log_scale);
    dqo = (adq ^ sgn) - sgn;
  }
  qcoeff[rc] = qo;
  dqcoeff[rc] = dqo;
}

hipError_t quantize_plain(const int32_t* coeff, int32_t* qcoeff, int32_t* dqcoeff, int n,
                          const PlainQuant& p, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_quantize_plain, dim3((n + 255) / 256), dim3(256), 0, st, coeff, qcoeff,
                     dqcoeff, n, p);
  return hipGetLastError();
}

// Plain-form high-bit-depth quantizers (quantize_fp_hbd / quantize_b_hbd,
// rust/crates/svtav1-encoder/src/quant.rs:335 / :391): the quantizer row as
// built (log_scale applied here), no INT16 clamp, levels written in raster
// order through the scan, eob = 1 + the last scan index with a level.  One
// workgroup walks the scan; the eob is a workgroup max.
__global__ __launch_bounds__(256) void k_quantize_hbd_plain(const int32_t* __restrict__ coeff,
                                                            const uint16_t* __restrict__ scan, int n, QuantRow t,
                                                            int ls, int kind, int32_t* __restrict__ qcoeff,
                                                            int32_t* __restrict__ dqcoeff, int32_t* __restrict__ eob) {
  __shared__ int32_t best[256];
  const int half = (1 << ls) >> 1;
  int32_t last = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int rc = scan[i], iz = rc != 0;
    const int32_t c = coeff[rc];
    const int32_t sgn = c >> 31;
    const int64_t a = (int64_t)((c ^ sgn) - sgn);
    int32_t lv = 0;
    if (kind == 0) {
      if ((a << (1 + ls)) >= t.dequant[iz])
        lv = (int32_t)(((a + ((t.round_fp[iz] + half) >> ls)) * t.quant_fp[iz]) >> (16 - ls));
    } else if (a >= ((t.zbin[iz] + half) >> ls)) {
      const int64_t x = a + ((t.round[iz] + half) >> ls);
      lv = (int32_t)(((((x * t.quant[iz]) >> 16) + x) * t.quant_shift[iz]) >> (16 - ls));
    }
    int32_t qo = 0, dqo = 0;
    if (lv != 0) {
      qo = (lv ^ sgn) - sgn;
      const int32_t adq = (int32_t)(((int64_t)lv * t.dequant[iz]) >> ls);
      dqo = (adq ^ sgn) - sgn;
      last = i + 1;
    }
    qcoeff[rc] = qo;
    dqcoeff[rc] = dqo;
  }
  best[threadIdx.x] = last;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) best[threadIdx.x] = max(best[threadIdx.x], best[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *eob = best[0];
}

hipError_t quantize_hbd_plain(const int32_t* coeff, const uint16_t* scan, int n, const QuantRow& row, int ls,
                              int kind, int32_t* qcoeff, int32_t* dqcoeff, int32_t* eob, hipStream_t st) {
  hipLaunchKernelGGL(k_quantize_hbd_plain, dim3(1), dim3(256), 0, st, coeff, scan, n, row, ls, kind, qcoeff, dqcoeff,
                     eob);
  return hipGetLastError();
}

// Plain-form CDEF distortion (compute_cdef_dist_8bit / _16bit,
// rust/crates/svtav1-dsp/src/cdef.rs:1192 / :1160): the SSE of packed
// filtered blocks against their source blocks (both gathered tight, bw x bh
// per block, block-major), rows stepped by sub, >> 2 * coeff_shift.
__global__ __launch_bounds__(256) void k_cdef_dist_plain(const uint16_t* __restrict__ src,
                                                         const uint16_t* __restrict__ packed, int n_samples, int bw,
                                                         int sub, int coeff_shift, uint64_t* __restrict__ out) {
  __shared__ uint64_t part[256];
  uint64_t acc = 0;
  for (int k = threadIdx.x; k < n_samples; k += 256) {
    const int row = k / bw;  // row index over the stacked blocks; every block starts at a multiple of sub
    if (row % sub) continue;
    const int32_t e = (int32_t)src[k] - (int32_t)packed[k];
    acc += (uint64_t)(uint32_t)(e * e);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0] >> (2 * coeff_shift);
}

hipError_t cdef_dist_plain(const uint16_t* src, const uint16_t* packed, int n_samples, int bw, int sub,
                           int coeff_shift, uint64_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_cdef_dist_plain, dim3(1), dim3(256), 0, st, src, packed, n_samples, bw, sub, coeff_shift, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SSE: one wavefront per job, lanes stride the job, DPP/shuffle reduction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_sse_i32(const int32_t* __restrict__ a,
                                                 const int32_t* __restrict__ b,
                                                 uint64_t* __restrict__ out, int n, int n_jobs) {
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= n_jobs) return;
  const int32_t* pa = a + (size_t)wave * n;
  const int32_t* pb = b ? b + (size_t)wave * n : nullptr;
  uint64_t acc = 0;
  for (int i = lane; i < n; i += 64) {
    const int64_t e = (int64_t)pa[i] - (pb ? (int64_t)pb[i] : 0);
    acc += (uint64_t)e * (uint64_t)e;
  }
  acc = wave_sum_u64(acc);
  if (lane == 0) out[wave] = acc;
}

hipError_t sse_i32_batch(const int32_t* a, const int32_t* b, uint64_t* out, int n_per_job,
                         int n_jobs, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
  const int grid = (n_jobs + 3) / 4;
  hipLaunchKernelGGL(k_sse_i32, dim3(grid), dim3(256), 0, st, a, b, out, n_per_job, n_jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Hadamard / SATD.  Each lane owns one 8x8 (or 4x4) sub-tile in registers;
// 16x16 and 32x32 combine their quadrants through LDS with the 16-bit wrap /
// saturation rules of the kernels the reference reproduces.
// ---------------------------------------------------------------------------
// Tile sources for k_satd: tile(i, &a) binds reader `a` to tile i and says
// whether the tile is live; emit(i, satd) stores its SATD.
struct PackedTiles {
  using Reader = PackedRes;
  const int16_t* res;
  int32_t* satd;
  int n, nn;  // tiles, coefficients per tile
  __device__ __forceinline__ bool tile(int i, int st, PackedRes* a) const {
    *a = {res + (size_t)i * nn, st};
    return i < n;
  }
  __device__ __forceinline__ void emit(int i, int32_t v) const { satd[i] = v; }
};

// Square tiles of (unit, prediction) jobs: tile i is tile k = i % tpj (raster
// order) of job i / tpj; the job's SATD is the sum over its tiles, added
// into out[] (0, or kSatdInvalid for a job outside the plane or the
// prediction buffer, as k_satd_init leaves it).
template <typename Pix>
struct UnitTiles {
  using Reader = DiffRes<Pix>;
  const Pix* src;
  int stride, pw, ph;
  const uint16_t* pred;
  long pred_len;
  const TxUnitJob* jobs;
  int w, h, t, tpr, tpj, n;  // n = tiles in all
  uint32_t* out;
  __device__ __forceinline__ bool tile(int i, int, DiffRes<Pix>* a) const {
    if (i >= n) return false;
    const int j = i / tpj, k = i - j * tpj;
    const TxUnitJob jb = jobs[j];
    if (!unit_inside(jb, w, h, pw, ph, pred_len)) return false;
    const int kr = k / tpr, r0 = kr * t, c0 = (k - kr * tpr) * t;
    *a = {src + (long)(jb.y + r0) * stride + jb.x + c0, pred + jb.pred_off + r0 * w + c0, stride, w};
    return true;
  }
  __device__ __forceinline__ void emit(int i, int32_t v) const {
    if (tpj == 1) out[i] = (uint32_t)v;
    else atomicAdd(out + i / tpj, (uint32_t)v);
  }
};

template <int N, class Src>
__global__ __launch_bounds__(256) void k_satd(Src tiles, int32_t* __restrict__ coeff_out) {
  constexpr int TPJ = N <= 8 ? 1 : (N / 8) * (N / 8);  // lanes per tile
  constexpr int JPW = 256 / TPJ;
  constexpr int NN = N * N;
  const int jl = threadIdx.x / TPJ, s = threadIdx.x - jl * TPJ;
  const int job = blockIdx.x * JPW + jl;
  typename Src::Reader src{};
  const bool active = tiles.tile(job, N, &src);
  if constexpr (N <= 8) {
    if (!active) return;
    int32_t c[64];
    if constexpr (N == 4) had4_tile(src, c);
    else had8_tile(src, c);
    int32_t sum = 0;
#pragma unroll
    for (int i = 0; i < NN; i++) {
      sum += c[i] < 0 ? -c[i] : c[i];
      if (coeff_out) coeff_out[(size_t)job * NN + i] = c[i];
    }
    tiles.emit(job, sum);
  } else {
    __shared__ int32_t lds[JPW * NN];
    __shared__ int32_t part[256];
    int32_t* L = lds + jl * NN;
    // sub-tile s: for N == 16 quadrant q = s; for N == 32, 16x16 quadrant Q = s / 4
    // and 8x8 quadrant q = s % 4 inside it (coefficient layout of the reference).
    int base_r, base_c, dst;
    if constexpr (N == 16) {
      base_r = (s >> 1) * 8; base_c = (s & 1) * 8; dst = s * 64;
    } else {
      const int Q = s >> 2, q = s & 3;
      base_r = (Q >> 1) * 16 + (q >> 1) * 8;
      base_c = (Q & 1) * 16 + (q & 1) * 8;
      dst = Q * 256 + q * 64;
    }
    if (active) {
      int32_t c[64];
      had8_tile(src.at(base_r, base_c), c);
#pragma unroll
      for (int i = 0; i < 64; i++) L[dst + i] = c[i];
    }
    __syncthreads();
    // 16x16 combine (int16 wrap, >> 1): per 16x16 group g, positions 0..63.
    constexpr int G16 = N == 16 ? 1 : 4;
    constexpr int PER = 64 * G16 / TPJ;  // positions per lane
    if (active) {
      for (int k = 0; k < PER; k++) {
        const int p = s * PER + k;
        const int g = p >> 6, i = p & 63;
        int32_t* B = L + g * 256;
        const int16_t a0 = (int16_t)B[i], a1 = (int16_t)B[i + 64];
        const int16_t a2 = (int16_t)B[i + 128], a3 = (int16_t)B[i + 192];
        const int16_t b0 = (int16_t)((int16_t)(a0 + a1) >> 1), b1 = (int16_t)((int16_t)(a0 - a1) >> 1);
        const int16_t b2 = (int16_t)((int16_t)(a2 + a3) >> 1), b3 = (int16_t)((int16_t)(a2 - a3) >> 1);
        B[i] = (int16_t)(b0 + b2);
        B[i + 64] = (int16_t)(b1 + b3);
        B[i + 128] = (int16_t)(b0 - b2);
        B[i + 192] = (int16_t)(b1 - b3);
      }
    }
    if constexpr (N == 32) {
      __syncthreads();
      if (active) {
        for (int k = 0; k < 16; k++) {
          const int i = s * 16 + k;
          const int32_t a0 = (int16_t)L[i], a1 = (int16_t)L[i + 256];
          const int32_t a2 = (int16_t)L[i + 512], a3 = (int16_t)L[i + 768];
          const int16_t b0 = sat16((a0 + a1) >> 2), b1 = sat16((a0 - a1) >> 2);
          const int16_t b2 = sat16((a2 + a3) >> 2), b3 = sat16((a2 - a3) >> 2);
          L[i] = (int16_t)(b0 + b2);
          L[i + 256] = (int16_t)(b1 + b3);
          L[i + 512] = (int16_t)(b0 - b2);
          L[i + 768] = (int16_t)(b1 - b3);
        }
      }
    }
    __syncthreads();
    int32_t sum = 0;
    if (active) {
      for (int i = s; i < NN; i += TPJ) {
        const int32_t v = L[i];
        sum += v < 0 ? -v : v;
        if (coeff_out) coeff_out[(size_t)job * NN + i] = v;
      }
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    if (active && s == 0) {
      int32_t tot = 0;
      for (int k = 0; k < TPJ; k++) tot += part[jl * TPJ + k];
      tiles.emit(job, tot);
    }
  }
}

hipError_t satd_batch(const int16_t* res, int32_t* coeff, int32_t* satd, int n, int n_jobs,
                      hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
  auto grid_for = [&](int tpj) { return (n_jobs + (256 / tpj) - 1) / (256 / tpj); };
  const PackedTiles src{res, satd, n_jobs, n * n};
  switch (n) {
    case 4: hipLaunchKernelGGL((k_satd<4, PackedTiles>), dim3(grid_for(1)), dim3(256), 0, st, src, coeff); break;
    case 8: hipLaunchKernelGGL((k_satd<8, PackedTiles>), dim3(grid_for(1)), dim3(256), 0, st, src, coeff); break;
    case 16: hipLaunchKernelGGL((k_satd<16, PackedTiles>), dim3(grid_for(4)), dim3(256), 0, st, src, coeff); break;
    case 32: hipLaunchKernelGGL((k_satd<32, PackedTiles>), dim3(grid_for(16)), dim3(256), 0, st, src, coeff); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Per job: 0 (its tiles add into it) or kSatdInvalid.
__global__ __launch_bounds__(256) void k_satd_init(const TxUnitJob* __restrict__ jobs, int n_jobs, int w, int h,
                                                   int pw, int ph, long pred_len, uint32_t* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n_jobs) return;
  out[j] = unit_inside(jobs[j], w, h, pw, ph, pred_len) ? 0u : kSatdInvalid;
}

template <typename Pix>
static hipError_t launch_unit_satd(const UnitTiles<Pix>& u, hipStream_t st) {
  auto grid_for = [&](int tpj) { return (u.n + (256 / tpj) - 1) / (256 / tpj); };
  switch (u.t) {
    case 4: hipLaunchKernelGGL((k_satd<4, UnitTiles<Pix>>), dim3(grid_for(1)), dim3(256), 0, st, u, nullptr); break;
    case 8: hipLaunchKernelGGL((k_satd<8, UnitTiles<Pix>>), dim3(grid_for(1)), dim3(256), 0, st, u, nullptr); break;
    case 16:
      hipLaunchKernelGGL((k_satd<16, UnitTiles<Pix>>), dim3(grid_for(4)), dim3(256), 0, st, u, nullptr);
      break;
    case 32:
      hipLaunchKernelGGL((k_satd<32, UnitTiles<Pix>>), dim3(grid_for(16)), dim3(256), 0, st, u, nullptr);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t pred_satd_batch(const void* src, int stride, int pw, int ph, int bd, int w, int h, const uint16_t* pred,
                           long pred_len, const TxUnitJob* jobs, int n_jobs, uint32_t* satd, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
  int t = w < h ? w : h;
  t = t > 32 ? 32 : t;
  const int tpr = w / t, tpj = tpr * (h / t);
  if ((long)n_jobs * tpj > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_satd_init, dim3((n_jobs + 255) / 256), dim3(256), 0, st, jobs, n_jobs, w, h, pw, ph, pred_len,
                     satd);
  if (bd == 8)
    return launch_unit_satd(UnitTiles<uint8_t>{(const uint8_t*)src, stride, pw, ph, pred, pred_len, jobs, w, h, t,
                                               tpr, tpj, n_jobs * tpj, satd},
                            st);
  return launch_unit_satd(UnitTiles<uint16_t>{(const uint16_t*)src, stride, pw, ph, pred, pred_len, jobs, w, h, t,
                                              tpr, tpj, n_jobs * tpj, satd},
                          st);
}

// ---------------------------------------------------------------------------
// Plain form: residual_i32 / recon_add_clamp / variance::sse (and their 16-bit
// twins) over one staged block of n packed samples.
// ---------------------------------------------------------------------------
template <typename Pix>
__global__ __launch_bounds__(256) void k_plain_residual(const Pix* __restrict__ src, const Pix* __restrict__ pred,
                                                        int n, int32_t* __restrict__ out) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) out[i] = (int32_t)src[i] - (int32_t)pred[i];
}

template <typename Pix>
__global__ __launch_bounds__(256) void k_plain_recon(const Pix* __restrict__ pred, const int32_t* __restrict__ inv,
                                                     int n, int32_t maxv, Pix* __restrict__ out) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int32_t v = (int32_t)pred[i] + inv[i];
    out[i] = (Pix)(v < 0 ? 0 : (v > maxv ? maxv : v));
  }
}

template <typename Pix>
__global__ __launch_bounds__(256) void k_plain_sse(const Pix* __restrict__ a, const Pix* __restrict__ b, int n,
                                                   unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int32_t d = (int32_t)a[i] - (int32_t)b[i];
    acc += (uint64_t)((int64_t)d * d);
  }
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)acc);
}

static unsigned plain_grid(int n) {
  const int g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 1024 ? 1024 : g));
}

hipError_t plain_residual(const void* src, const void* pred, int n, int hbd, int32_t* out, hipStream_t st) {
  if (hbd)
    hipLaunchKernelGGL((k_plain_residual<uint16_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint16_t*)src,
                       (const uint16_t*)pred, n, out);
  else
    hipLaunchKernelGGL((k_plain_residual<uint8_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint8_t*)src,
                       (const uint8_t*)pred, n, out);
  return hipGetLastError();
}

hipError_t plain_recon(const void* pred, const int32_t* inv, int n, int bd, int hbd, void* out, hipStream_t st) {
  const int32_t maxv = (1 << bd) - 1;
  if (hbd)
    hipLaunchKernelGGL((k_plain_recon<uint16_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint16_t*)pred, inv,
                       n, maxv, (uint16_t*)out);
  else
    hipLaunchKernelGGL((k_plain_recon<uint8_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint8_t*)pred, inv, n,
                       maxv, (uint8_t*)out);
  return hipGetLastError();
}

// out must be zeroed on the stream first (a fill kernel, not a memset node)
__global__ void k_zero_u64_one(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0;
}

hipError_t plain_sse(const void* a, const void* b, int n, int hbd, uint64_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64_one, dim3(1), dim3(64), 0, st, (unsigned long long*)out);
  if (hbd)
    hipLaunchKernelGGL((k_plain_sse<uint16_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint16_t*)a,
                       (const uint16_t*)b, n, (unsigned long long*)out);
  else
    hipLaunchKernelGGL((k_plain_sse<uint8_t>), dim3(plain_grid(n)), dim3(256), 0, st, (const uint8_t*)a,
                       (const uint8_t*)b, n, (unsigned long long*)out);
  return hipGetLastError();
}

}  // namespace zav1
