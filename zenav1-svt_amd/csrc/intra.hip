// Batched AV1 intra prediction for gfx950 (8- and 10-bit planes).
//
// One 64-lane wavefront per transform unit: the wave assembles the above /
// left reference edges in LDS straight from the reconstruction plane
// (availability counts come from the host), applies the corner / edge
// filters and 2x upsampling, then predicts every sample of the block with
// lanes striding the w x h output.  Filter intra runs its 4x2 sub-blocks as
// an anti-diagonal wavefront (each sub-block needs its left, upper and
// upper-left neighbours).
//
// Semantics: rust/crates/svtav1-encoder/src/leaf_funnel/predict.rs:49
// (predict_unit routing), rust/crates/svtav1-encoder/src/partition.rs:250
// (neighbour fill), rust/crates/svtav1-encoder/src/intra_edge.rs:910
// (directional edge assembly + filters), rust/crates/svtav1-dsp/src/
// intra_pred.rs:18-1180 and hbd.rs:169-935 (predictors, CfL).
#include "av1_intra_tables.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

namespace {
constexpr int kO = 16;      // edge origin (the C `above_data + 16`)
constexpr int kEdge = 160;  // edge buffer length

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
__device__ __forceinline__ int dr_dx(int a) {
  return (a > 0 && a < 90) ? DR_INTRA_DERIVATIVE[a] : ((a > 90 && a < 180) ? DR_INTRA_DERIVATIVE[180 - a] : 1);
}
__device__ __forceinline__ int dr_dy(int a) {
  return (a > 90 && a < 180) ? DR_INTRA_DERIVATIVE[a - 90] : ((a > 180 && a < 270) ? DR_INTRA_DERIVATIVE[270 - a] : 1);
}
__device__ __forceinline__ int edge_strength(int bs0, int bs1, int delta, int ft) {
  const int d = delta < 0 ? -delta : delta, wh = bs0 + bs1;
  int s = 0;
  if (ft == 0) {
    if (wh <= 8) { if (d >= 56) s = 1; }
    else if (wh <= 16) { if (d >= 40) s = 1; }
    else if (wh <= 24) { if (d >= 8) s = 1; if (d >= 16) s = 2; if (d >= 32) s = 3; }
    else if (wh <= 32) { if (d >= 1) s = 1; if (d >= 4) s = 2; if (d >= 32) s = 3; }
    else if (d >= 1) s = 3;
  } else {
    if (wh <= 8) { if (d >= 40) s = 1; if (d >= 64) s = 2; }
    else if (wh <= 16) { if (d >= 20) s = 1; if (d >= 48) s = 2; }
    else if (wh <= 24) { if (d >= 4) s = 3; }
    else if (d >= 1) s = 3;
  }
  return s;
}
__device__ __forceinline__ bool use_upsample(int bs0, int bs1, int delta, int ft) {
  const int d = delta < 0 ? -delta : delta, wh = bs0 + bs1;
  if (d <= 0 || d >= 40) return false;
  return ft ? (wh <= 8) : (wh <= 16);
}
__device__ __forceinline__ const uint8_t* sm_w(int n) {
  return SM_WEIGHTS_ALL + (n == 4 ? 0 : n == 8 ? 4 : n == 16 ? 12 : n == 32 ? 28 : 60);
}

// 5-tap smoothing of e[start .. start+sz) in place (tmp = scratch copy).
__device__ void filter_edge_wave(int32_t* e, int32_t* tmp, int start, int sz, int strength, int lane) {
  if (strength == 0) return;  // wave-uniform
  for (int i = lane; i < sz; i += 64) tmp[i] = e[start + i];
  __syncthreads();
  const int k0 = strength == 1 ? 0 : (strength == 2 ? 0 : 2);
  const int k1 = strength == 1 ? 4 : (strength == 2 ? 5 : 4);
  const int k2 = strength == 1 ? 8 : (strength == 2 ? 6 : 4);
  for (int i = 1 + lane; i < sz; i += 64) {
    auto at = [&](int k) { return tmp[clampi(k, 0, sz - 1)]; };
    const int32_t s = at(i - 2) * k0 + at(i - 1) * k1 + at(i) * k2 + at(i + 1) * k1 + at(i + 2) * k0;
    e[start + i] = (s + 8) >> 4;
  }
  __syncthreads();
}

__device__ void upsample_wave(int32_t* p, int32_t* tmp, int sz, int maxv, int lane) {
  // tmp[0..sz+3) = {p[o-1], p[o-1], p[o..o+sz), p[o+sz-1]}
  for (int i = lane; i < sz + 3; i += 64)
    tmp[i] = (i < 2) ? p[kO - 1] : (i < sz + 2 ? p[kO + i - 2] : p[kO + sz - 1]);
  __syncthreads();
  if (lane == 0) p[kO - 2] = tmp[0];
  for (int i = lane; i < sz; i += 64) {
    const int32_t s = -tmp[i] + 9 * tmp[i + 1] + 9 * tmp[i + 2] - tmp[i + 3];
    p[kO + 2 * i - 1] = clampi((s + 8) >> 4, 0, maxv);
    p[kO + 2 * i] = tmp[i + 2];
  }
  __syncthreads();
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
}  // namespace

template <typename Pix>
__global__ __launch_bounds__(64) void k_intra_pred(const Pix* __restrict__ plane, int stride, int bd,
                                                   const IntraJob* __restrict__ jobs, int n_jobs,
                                                   uint16_t* __restrict__ out) {
  __shared__ int32_t above[kEdge], left[kEdge], tmp[kEdge];
  __shared__ int32_t fbuf[33 * 33];
  const int lane = threadIdx.x;
  const int ji = blockIdx.x;
  if (ji >= n_jobs) return;  // whole workgroup exits together
  const IntraJob j = jobs[ji];
  const int w = j.w, h = j.h, n = w * h;
  const int base = 128 << (bd - 8), maxv = (1 << bd) - 1;
  uint16_t* dst = out + j.out_off;
  auto P = [&](int x, int y) -> int32_t { return (int32_t)plane[(long)y * stride + x]; };

  const bool directional = (j.mode >= 3 && j.mode <= 8) || ((j.mode == 1 || j.mode == 2) && j.delta != 0);
  if (directional) {
    const int pa = MODE_TO_ANGLE[j.mode] + 3 * j.delta;
    const bool need_above = pa < 180, need_left = pa > 90;
    const bool need_right = pa < 90, need_bottom = pa > 180;
    const int n_top = j.n_top_px, n_tr = j.n_topright_px, n_left = j.n_left_px, n_bl = j.n_bottomleft_px;
    if ((!need_above && n_left == 0) || (!need_left && n_top == 0)) {
      const int32_t v = need_left ? (n_top > 0 ? P(j.x, j.y - 1) : base + 1)
                                  : (n_left > 0 ? P(j.x - 1, j.y) : base - 1);
      for (int i = lane; i < n; i += 64) dst[i] = (uint16_t)v;
      return;
    }
    for (int i = lane; i < kEdge; i += 64) {
      above[i] = base - 1;
      left[i] = base + 1;
    }
    __syncthreads();
    if (need_left) {
      const int num_left = h + (need_bottom ? w : 0);
      if (n_left > 0) {
        const int real = (need_bottom && n_bl > 0) ? h + n_bl : n_left;
        for (int i = lane; i < num_left; i += 64) left[kO + i] = P(j.x - 1, j.y + (i < real ? i : real - 1));
      } else if (n_top > 0) {
        const int32_t v = P(j.x, j.y - 1);
        for (int i = lane; i < num_left; i += 64) left[kO + i] = v;
      }
    }
    if (need_above) {
      const int num_top = w + (need_right ? h : 0);
      if (n_top > 0) {
        const int real = (need_right && n_tr > 0) ? w + n_tr : n_top;
        for (int i = lane; i < num_top; i += 64) above[kO + i] = P(j.x + (i < real ? i : real - 1), j.y - 1);
      } else if (n_left > 0) {
        const int32_t v = P(j.x - 1, j.y);
        for (int i = lane; i < num_top; i += 64) above[kO + i] = v;
      }
    }
    if (lane == 0) {
      int32_t tl;
      if (n_top > 0 && n_left > 0) tl = P(j.x - 1, j.y - 1);
      else if (n_top > 0) tl = P(j.x, j.y - 1);
      else if (n_left > 0) tl = P(j.x - 1, j.y);
      else tl = base;
      above[kO - 1] = tl;
      left[kO - 1] = tl;
    }
    __syncthreads();
    bool up_a = false, up_l = false;
    if (j.edge_filter) {
      if (pa != 90 && pa != 180) {
        if (need_above && need_left && (w + h >= 24)) {
          if (lane == 0) {
            const int32_t s = (left[kO] * 5 + above[kO - 1] * 6 + above[kO] * 5 + 8) >> 4;
            above[kO - 1] = s;
            left[kO - 1] = s;
          }
          __syncthreads();
        }
        if (need_above && n_top > 0)
          filter_edge_wave(above, tmp, kO - 1, n_top + 1 + (need_right ? h : 0),
                           edge_strength(w, h, pa - 90, j.filt_type), lane);
        if (need_left && n_left > 0)
          filter_edge_wave(left, tmp, kO - 1, n_left + 1 + (need_bottom ? w : 0),
                           edge_strength(h, w, pa - 180, j.filt_type), lane);
      }
      up_a = use_upsample(w, h, pa - 90, j.filt_type);
      if (need_above && up_a) upsample_wave(above, tmp, w + (need_right ? h : 0), maxv, lane);
      up_l = use_upsample(h, w, pa - 180, j.filt_type);
      if (need_left && up_l) upsample_wave(left, tmp, h + (need_bottom ? w : 0), maxv, lane);
    }
    const int dx = dr_dx(pa), dy = dr_dy(pa);
    const int ua = up_a ? 1 : 0, ul = up_l ? 1 : 0;
    for (int i = lane; i < n; i += 64) {
      const int r = i / w, c = i - r * w;
      int32_t v;
      if (pa < 90) {  // zone 1: along the above edge
        const int max_base = (w + h - 1) << ua;
        const int x = dx * (r + 1);
        const int b0 = x >> (6 - ua), shift = ((x << ua) & 0x3F) >> 1;
        const int b = b0 + (c << ua);
        if (b0 >= max_base || b >= max_base) v = above[kO + max_base];
        else v = clampi((above[kO + b] * (32 - shift) + above[kO + b + 1] * shift + 16) >> 5, 0, maxv);
      } else if (pa < 180) {  // zone 2
        const int x = (c << 6) - (r + 1) * dx, b = x >> (6 - ua);
        int32_t s;
        if (b >= -(1 << ua)) {
          const int sh = ((x * (1 << ua)) & 0x3F) >> 1;
          s = above[kO + b] * (32 - sh) + above[kO + b + 1] * sh;
        } else {
          const int y = (r << 6) - (c + 1) * dy, b2 = y >> (6 - ul);
          const int sh = ((y * (1 << ul)) & 0x3F) >> 1;
          s = left[kO + b2] * (32 - sh) + left[kO + b2 + 1] * sh;
        }
        v = clampi((s + 16) >> 5, 0, maxv);
      } else {  // zone 3: along the left edge
        const int max_base = (w + h - 1) << ul;
        const int y = dy * (c + 1);
        const int b = (y >> (6 - ul)) + (r << ul), shift = ((y << ul) & 0x3F) >> 1;
        if (b >= max_base) v = left[kO + max_base];
        else v = clampi((left[kO + b] * (32 - shift) + left[kO + b + 1] * shift + 16) >> 5, 0, maxv);
      }
      dst[i] = (uint16_t)v;
    }
    return;
  }

  // ---- non-directional / filter intra: neighbour fill ----
  const bool has_above = j.n_top_px > 0, has_left = j.n_left_px > 0;
  const int32_t a0 = has_above ? P(j.x, j.y - 1) : 0;
  const int32_t l0 = has_left ? P(j.x - 1, j.y) : 0;
  for (int i = lane; i < w; i += 64)
    above[i] = has_above ? P(j.x + (i < j.n_top_px ? i : j.n_top_px - 1), j.y - 1) : (has_left ? l0 : base - 1);
  for (int i = lane; i < h; i += 64)
    left[i] = has_left ? P(j.x - 1, j.y + (i < j.n_left_px ? i : j.n_left_px - 1)) : (has_above ? a0 : base + 1);
  const int32_t tl = (has_above && has_left) ? P(j.x - 1, j.y - 1) : (has_above ? a0 : (has_left ? l0 : base));
  __syncthreads();

  if (j.fi_mode < 5) {
    // 33x33 staging buffer; sub-block (i, k) of 4x2 samples at rows 1+2i, cols 1+4k.
    for (int i = lane; i < 33 * 33; i += 64) fbuf[i] = 0;
    __syncthreads();
    if (lane == 0) fbuf[0] = tl;
    for (int c = lane; c < w; c += 64) fbuf[c + 1] = above[c];
    for (int r = lane; r < h; r += 64) fbuf[(r + 1) * 33] = left[r];
    __syncthreads();
    const int8_t* taps = FILTER_INTRA_TAPS + j.fi_mode * 64;
    const int sbr = h / 2, sbc = w / 4;
    for (int d = 0; d < sbr + sbc - 1; d++) {
      // lanes: (sub-block on diagonal d) x (8 output samples)
      const int sb = lane >> 3, k = lane & 7;
      for (int s = sb; s < sbr; s += 8) {
        const int bc = d - s;
        if (bc < 0 || bc >= sbc) continue;
        const int r = 1 + 2 * s, c = 1 + 4 * bc;
        const int32_t p0 = fbuf[(r - 1) * 33 + c - 1], p1 = fbuf[(r - 1) * 33 + c], p2 = fbuf[(r - 1) * 33 + c + 1];
        const int32_t p3 = fbuf[(r - 1) * 33 + c + 2], p4 = fbuf[(r - 1) * 33 + c + 3];
        const int32_t p5 = fbuf[r * 33 + c - 1], p6 = fbuf[(r + 1) * 33 + c - 1];
        const int8_t* t = taps + k * 8;
        const int32_t v = t[0] * p0 + t[1] * p1 + t[2] * p2 + t[3] * p3 + t[4] * p4 + t[5] * p5 + t[6] * p6;
        const int32_t rv = v < 0 ? -((-v + 8) >> 4) : ((v + 8) >> 4);
        fbuf[(r + (k >> 2)) * 33 + c + (k & 3)] = clampi(rv, 0, maxv);
      }
      __syncthreads();
    }
    for (int i = lane; i < n; i += 64) {
      const int r = i / w, c = i - r * w;
      dst[i] = (uint16_t)fbuf[(r + 1) * 33 + c + 1];
    }
    return;
  }

  int32_t dc = 0;
  if (j.mode == 0) {
    int32_t s = 0;
    if (has_above)
      for (int i = lane; i < w; i += 64) s += above[i];
    if (has_left)
      for (int i = lane; i < h; i += 64) s += left[i];
    s = wave_sum(s);
    const uint32_t us = (uint32_t)s;
    if (has_above && has_left) dc = (int32_t)((us + (w + h) / 2) / (w + h));
    else if (has_above) dc = (int32_t)((us + w / 2) / w);
    else if (has_left) dc = (int32_t)((us + h / 2) / h);
    else dc = base;
  }
  const uint8_t *wh = sm_w(h), *ww = sm_w(w);
  const bool hbd = bd > 8;
  for (int i = lane; i < n; i += 64) {
    const int r = i / w, c = i - r * w;
    int32_t v;
    switch (j.mode) {
      case 0: v = dc; break;
      case 1: v = above[c]; break;
      case 2: v = left[r]; break;
      case 9: {
        const uint32_t p = (uint32_t)(wh[r] * above[c] + (256 - wh[r]) * left[h - 1] + ww[c] * left[r] +
                                      (256 - ww[c]) * above[w - 1] + 256) / 512;
        v = p > (uint32_t)maxv ? maxv : (int32_t)p;
        break;
      }
      case 10: {
        const uint32_t p = (uint32_t)(wh[r] * above[c] + (256 - wh[r]) * left[h - 1] + 128) / 256;
        v = p > (uint32_t)maxv ? maxv : (int32_t)p;
        break;
      }
      case 11: {
        const uint32_t p = (uint32_t)(ww[c] * left[r] + (256 - ww[c]) * above[w - 1] + 128) / 256;
        v = p > (uint32_t)maxv ? maxv : (int32_t)p;
        break;
      }
      default: {  // 12: Paeth (8-bit top-first tie order, high bit depth left-first)
        const int32_t top = above[c], lf = left[r], b = top + lf - tl;
        const int32_t pt = abs(b - top), pl = abs(b - lf), ptl = abs(b - tl);
        if (!hbd) v = (pt <= pl && pt <= ptl) ? top : (pl <= ptl ? lf : tl);
        else v = (pl <= pt && pl <= ptl) ? lf : (pt <= ptl ? top : tl);
        break;
      }
    }
    dst[i] = (uint16_t)v;
  }
}

// CfL: one wave per chroma block.  luma: 2cw x 2ch reconstructed luma at
// (lx, ly); dc: chroma DC prediction (cw x ch at dc_off); out at out_off.
template <typename Pix>
__global__ __launch_bounds__(64) void k_cfl(const Pix* __restrict__ luma, int stride, int bd,
                                            const CflJob* __restrict__ jobs, int n_jobs,
                                            const uint16_t* __restrict__ dcbuf, uint16_t* __restrict__ out) {
  const int lane = threadIdx.x, ji = blockIdx.x;
  if (ji >= n_jobs) return;
  const CflJob j = jobs[ji];
  const int n = j.cw * j.ch, maxv = (1 << bd) - 1;
  int32_t s = 0;
  for (int i = lane; i < n; i += 64) {
    const int r = i / j.cw, c = i - r * j.cw;
    const Pix* p = luma + (long)(j.ly + 2 * r) * stride + j.lx + 2 * c;
    s += (int16_t)(((int32_t)p[0] + p[1] + p[stride] + p[stride + 1]) * 2);
  }
  s = wave_sum(s);
  int l2 = 0;
  while ((1 << l2) < n) l2++;
  const int32_t avg = (s + ((1 << l2) >> 1)) >> l2;
  for (int i = lane; i < n; i += 64) {
    const int r = i / j.cw, c = i - r * j.cw;
    const Pix* p = luma + (long)(j.ly + 2 * r) * stride + j.lx + 2 * c;
    const int16_t q3 = (int16_t)(((int32_t)p[0] + p[1] + p[stride] + p[stride + 1]) * 2);
    const int16_t ac = (int16_t)(q3 - avg);
    const int32_t sc = j.alpha_q3 * ac;
    const int32_t q0 = sc < 0 ? -((-sc + 32) >> 6) : ((sc + 32) >> 6);
    out[j.out_off + i] = (uint16_t)clampi(q0 + dcbuf[j.dc_off + i], 0, maxv);
  }
}

hipError_t intra_pred_batch(const void* plane, int stride, int bd, const IntraJob* jobs, int n_jobs, uint16_t* out,
                            hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
  if (bd == 8)
    hipLaunchKernelGGL((k_intra_pred<uint8_t>), dim3(n_jobs), dim3(64), 0, st, (const uint8_t*)plane, stride, bd,
                       jobs, n_jobs, out);
  else
    hipLaunchKernelGGL((k_intra_pred<uint16_t>), dim3(n_jobs), dim3(64), 0, st, (const uint16_t*)plane, stride,
                       bd, jobs, n_jobs, out);
  return hipGetLastError();
}

hipError_t cfl_batch(const void* luma, int stride, int bd, const CflJob* jobs, int n_jobs, const uint16_t* dc,
                     uint16_t* out, hipStream_t st) {
  if (n_jobs <= 0) return hipSuccess;
  if (bd == 8)
    hipLaunchKernelGGL((k_cfl<uint8_t>), dim3(n_jobs), dim3(64), 0, st, (const uint8_t*)luma, stride, bd, jobs,
                       n_jobs, dc, out);
  else
    hipLaunchKernelGGL((k_cfl<uint16_t>), dim3(n_jobs), dim3(64), 0, st, (const uint16_t*)luma, stride, bd, jobs,
                       n_jobs, dc, out);
  return hipGetLastError();
}

}  // namespace zav1
