// Mode-decision funnel selection for gfx950: after the MDS0 SATD of every
// (unit, candidate prediction) (pred_satd_batch, beside the Hadamard kernels),
// the per-unit top-K selection that feeds the transform-unit stage, on the
// device, so no cost table crosses to the host between the stages.
//
//   fast cost      rdcost(lambda, rate, satd << 4)
//                  rust/crates/svtav1-encoder/src/leaf_funnel/mod.rs:2181-2185,
//                  rdcost rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:1033
//   ordering       the swap-on-< exchange sort over the candidates in their
//                  generation order (not stable on equal costs),
//                  c_exchange_sort_by rust/crates/svtav1-encoder/src/leaf_funnel/rate_tables.rs:108,
//                  applied to the fast costs at mod.rs:2994
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

// One wave per unit (grid-stride over units).  Lane m computes candidate m's
// fast cost; the first k outer passes of the exchange sort (position p is
// final after pass p) run lane-parallel over the candidates in generation
// order (wave_exchange_first_k); the lanes then write the survivors'
// transform-unit jobs.
constexpr int kMaxCands = 64;
__global__ __launch_bounds__(256) void k_fast_select(const uint32_t* __restrict__ satd,
                                                     const uint32_t* __restrict__ rate, uint64_t lambda,
                                                     int n_units, int n_cands, int k,
                                                     const int32_t* __restrict__ unit_xy, int unit_stride, int w,
                                                     int h, TypeList types, int survivors_only,
                                                     uint8_t* __restrict__ sel,
                                                     uint64_t* __restrict__ sel_cost,
                                                     TxUnitJob* __restrict__ jobs_out) {
  __shared__ uint8_t ord[4][kMaxCands];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint8_t* O = ord[wave];
  for (int u = blockIdx.x * 4 + wave; u < n_units; u += gridDim.x * 4) {
    const long base = (long)u * n_cands;
    int idx = lane;
    uint64_t cost = ~0ull;
    if (lane < n_cands) {
      const uint32_t s = satd[base + lane];
      const uint64_t r = rate ? rate[base + lane] : 0u;
      cost = s == kSatdInvalid ? ~0ull : ((r * lambda + 256) >> 9) + ((uint64_t)s << 11);
    }
    wave_exchange_first_k(idx, cost, n_cands, k, lane, sel + (long)u * k, sel_cost ? sel_cost + (long)u * k : nullptr);
    O[lane] = (uint8_t)idx;
    wave_sync();
    const int ux = unit_xy[(long)u * unit_stride], uy = unit_xy[(long)u * unit_stride + 1];
    for (int t = lane; t < k * types.n; t += 64) {
      const int i = t / types.n, ti = t - i * types.n;
      TxUnitJob jb;
      jb.x = ux;
      jb.y = uy;
      jb.pred_off = (uint32_t)((survivors_only ? (long)u * k + i : base + O[i]) * (long)(w * h));
      jb.tx_type = types.t[ti];
      jb.crop_w = (uint8_t)w;
      jb.crop_h = (uint8_t)h;
      jb.pad = 0;
      jobs_out[((long)u * k + i) * types.n + ti] = jb;
    }
    wave_sync();
  }
}

hipError_t fast_select(const uint32_t* satd, const uint32_t* rate, uint64_t lambda, int n_units, int n_cands, int k,
                       const int32_t* unit_xy, int unit_stride, int w, int h, const TypeList& types,
                       int survivors_only, uint8_t* sel, uint64_t* sel_cost, TxUnitJob* jobs, hipStream_t st) {
  if (n_units <= 0) return hipSuccess;
  if (n_cands < 1 || n_cands > kMaxCands || k < 1 || k > n_cands) return hipErrorInvalidValue;
  const int grid = (n_units + 3) / 4 < 8192 ? (n_units + 3) / 4 : 8192;
  hipLaunchKernelGGL(k_fast_select, dim3(grid), dim3(256), 0, st, satd, rate, lambda, n_units,
                     n_cands, k, unit_xy, unit_stride, w, h, types, survivors_only, sel, sel_cost, jobs);
  return hipGetLastError();
}

}  // namespace zav1
