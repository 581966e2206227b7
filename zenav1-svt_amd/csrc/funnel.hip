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

// One lane per unit.  Position p of the exchange sort is final after its
// outer pass p, so the first k passes give the k survivors in the order the
// sort leaves them; the candidate order of the unit lives in LDS, transposed
// (position-major) so a wave's lanes touch consecutive bytes.
constexpr int kMaxCands = 64;
__global__ __launch_bounds__(256) void k_fast_select(const uint32_t* __restrict__ satd,
                                                     const uint32_t* __restrict__ rate, uint64_t lambda,
                                                     int n_units, int n_cands, int k,
                                                     const int32_t* __restrict__ unit_xy, int unit_stride, int w,
                                                     int h, TypeList types, int survivors_only,
                                                     uint8_t* __restrict__ sel,
                                                     uint64_t* __restrict__ sel_cost,
                                                     TxUnitJob* __restrict__ jobs_out) {
  __shared__ uint8_t ord[kMaxCands * 256];
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= n_units) return;
  const long base = (long)u * n_cands;
  auto cost = [&](int m) -> uint64_t {
    const uint32_t s = satd[base + m];
    if (s == kSatdInvalid) return ~0ull;
    const uint64_t r = rate ? rate[base + m] : 0u;
    return ((r * lambda + 256) >> 9) + ((uint64_t)s << 11);
  };
  for (int m = 0; m < n_cands; m++) ord[m * 256 + threadIdx.x] = (uint8_t)m;
  const int ux = unit_xy[(long)u * unit_stride], uy = unit_xy[(long)u * unit_stride + 1];
  for (int i = 0; i < k; i++) {
    int mi = ord[i * 256 + threadIdx.x];
    uint64_t ci = cost(mi);
    for (int j = i + 1; j < n_cands; j++) {
      const int mj = ord[j * 256 + threadIdx.x];
      const uint64_t cj = cost(mj);
      if (cj < ci) {  // swap positions i and j
        ord[j * 256 + threadIdx.x] = (uint8_t)mi;
        mi = mj;
        ci = cj;
      }
    }
    ord[i * 256 + threadIdx.x] = (uint8_t)mi;
    sel[(long)u * k + i] = (uint8_t)mi;
    if (sel_cost) sel_cost[(long)u * k + i] = ci;
    for (int ti = 0; ti < types.n; ti++) {
      TxUnitJob jb;
      jb.x = ux;
      jb.y = uy;
      jb.pred_off = (uint32_t)((survivors_only ? (long)u * k + i : base + mi) * (long)(w * h));
      jb.tx_type = types.t[ti];
      jb.crop_w = (uint8_t)(w > 255 ? 255 : w);
      jb.crop_h = (uint8_t)(h > 255 ? 255 : h);
      jb.pad = 0;
      jobs_out[((long)u * k + i) * types.n + ti] = jb;
    }
  }
}

hipError_t fast_select(const uint32_t* satd, const uint32_t* rate, uint64_t lambda, int n_units, int n_cands, int k,
                       const int32_t* unit_xy, int unit_stride, int w, int h, const TypeList& types,
                       int survivors_only, uint8_t* sel, uint64_t* sel_cost, TxUnitJob* jobs, hipStream_t st) {
  if (n_units <= 0) return hipSuccess;
  if (n_cands < 1 || n_cands > kMaxCands || k < 1 || k > n_cands) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fast_select, dim3((n_units + 255) / 256), dim3(256), 0, st, satd, rate, lambda, n_units,
                     n_cands, k, unit_xy, unit_stride, w, h, types, survivors_only, sel, sel_cost, jobs);
  return hipGetLastError();
}

}  // namespace zav1
