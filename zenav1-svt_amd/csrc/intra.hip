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
#include "hadamard.h"
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

// Barriers of the predictor: a workgroup of one wave uses the workgroup
// barrier; waves predicting different candidates of one workgroup order
// their own LDS traffic with a wave-level fence + barrier.
struct BlockSync {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
};
struct WaveSync {
  __device__ __forceinline__ void operator()() const { wave_sync(); }
};

// 5-tap smoothing of e[start .. start+sz) in place (tmp = scratch copy).
template <class SYNC>
__device__ void filter_edge_wave(int16_t* e, int16_t* tmp, int start, int sz, int strength, int lane, SYNC sync) {
  if (strength == 0) return;  // wave-uniform
  for (int i = lane; i < sz; i += 64) tmp[i] = e[start + i];
  sync();
  const int k0 = strength == 1 ? 0 : (strength == 2 ? 0 : 2);
  const int k1 = strength == 1 ? 4 : (strength == 2 ? 5 : 4);
  const int k2 = strength == 1 ? 8 : (strength == 2 ? 6 : 4);
  for (int i = 1 + lane; i < sz; i += 64) {
    auto at = [&](int k) { return tmp[clampi(k, 0, sz - 1)]; };
    const int32_t s = at(i - 2) * k0 + at(i - 1) * k1 + at(i) * k2 + at(i + 1) * k1 + at(i + 2) * k0;
    e[start + i] = (int16_t)((s + 8) >> 4);
  }
  sync();
}

template <class SYNC>
__device__ void upsample_wave(int16_t* p, int16_t* tmp, int sz, int maxv, int lane, SYNC sync) {
  // tmp[0..sz+3) = {p[o-1], p[o-1], p[o..o+sz), p[o+sz-1]}
  for (int i = lane; i < sz + 3; i += 64)
    tmp[i] = (i < 2) ? p[kO - 1] : (i < sz + 2 ? p[kO + i - 2] : p[kO + sz - 1]);
  sync();
  if (lane == 0) p[kO - 2] = tmp[0];
  for (int i = lane; i < sz; i += 64) {
    const int32_t s = -tmp[i] + 9 * tmp[i + 1] + 9 * tmp[i + 2] - tmp[i + 3];
    p[kO + 2 * i - 1] = (int16_t)clampi((s + 8) >> 4, 0, maxv);
    p[kO + 2 * i] = tmp[i + 2];
  }
  sync();
}

__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
}  // namespace

// Directional prediction from final edges (dr_predictor_edged,
// rust/crates/svtav1-dsp/src/intra_pred.rs:810): above[kO + i] / left[kO + i]
// are the (filtered, possibly upsampled: ua / ul) above row and left column,
// above[kO - 1] the corner; angle pa in (0, 270).  Lanes stride the w x h
// samples; out(i, v) receives sample i.
template <class OUT>
__device__ __forceinline__ void dr_predict_edges(int pa, const int16_t* above, const int16_t* left, int w, int h,
                                                 int ua, int ul, int maxv, const OUT& out, int lane) {
  const int dx = dr_dx(pa), dy = dr_dy(pa), n = w * h;
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
    } else if (pa == 90) {
      v = above[kO + c];
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
    } else if (pa == 180) {
      v = left[kO + r];
    } else {  // zone 3: along the left edge
      const int max_base = (w + h - 1) << ul;
      const int y = dy * (c + 1);
      const int b = (y >> (6 - ul)) + (r << ul), shift = ((y << ul) & 0x3F) >> 1;
      if (b >= max_base) v = left[kO + max_base];
      else v = clampi((left[kO + b] * (32 - shift) + left[kO + b + 1] * shift + 16) >> 5, 0, maxv);
    }
    out(i, v);
  }
}

// One unit's prediction by one wave: P(x, y) reads a reference sample of
// the plane (only the row above, the column left and the top-left corner
// of the unit, within the job's availability counts); out(i, v) receives
// sample i (raster) of the w x h prediction.  above / left / tmp hold kEdge
// samples, fbuf (h + 1) x 33 (filter intra only), all private to the wave.
// W, H > 0 fix the unit size at compile time (j.w, j.h otherwise).
template <int W = 0, int H = 0, class NB, class OUT, class SYNC>
__device__ void predict_wave(const IntraJob& j, int bd, const NB& P, const OUT& out, SYNC sync, int16_t* above,
                             int16_t* left, int16_t* tmp, int16_t* fbuf, int lane) {
  const int w = W > 0 ? W : j.w, h = H > 0 ? H : j.h, n = w * h;
  const int base = 128 << (bd - 8), maxv = (1 << bd) - 1;

  const bool directional = (j.mode >= 3 && j.mode <= 8) || ((j.mode == 1 || j.mode == 2) && j.delta != 0);
  if (directional) {
    const int pa = MODE_TO_ANGLE[j.mode] + 3 * j.delta;
    const bool need_above = pa < 180, need_left = pa > 90;
    const bool need_right = pa < 90, need_bottom = pa > 180;
    const int n_top = j.n_top_px, n_tr = j.n_topright_px, n_left = j.n_left_px, n_bl = j.n_bottomleft_px;
    if ((!need_above && n_left == 0) || (!need_left && n_top == 0)) {
      const int32_t v = need_left ? (n_top > 0 ? P(j.x, j.y - 1) : base + 1)
                                  : (n_left > 0 ? P(j.x - 1, j.y) : base - 1);
      for (int i = lane; i < n; i += 64) out(i, v);
      return;
    }
    for (int i = lane; i < kEdge; i += 64) {
      above[i] = base - 1;
      left[i] = base + 1;
    }
    sync();
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
    sync();
    bool up_a = false, up_l = false;
    if (j.edge_filter) {
      if (pa != 90 && pa != 180) {
        if (need_above && need_left && (w + h >= 24)) {
          if (lane == 0) {
            const int32_t s = (left[kO] * 5 + above[kO - 1] * 6 + above[kO] * 5 + 8) >> 4;
            above[kO - 1] = s;
            left[kO - 1] = s;
          }
          sync();
        }
        if (need_above && n_top > 0)
          filter_edge_wave(above, tmp, kO - 1, n_top + 1 + (need_right ? h : 0),
                           edge_strength(w, h, pa - 90, j.filt_type), lane, sync);
        if (need_left && n_left > 0)
          filter_edge_wave(left, tmp, kO - 1, n_left + 1 + (need_bottom ? w : 0),
                           edge_strength(h, w, pa - 180, j.filt_type), lane, sync);
      }
      up_a = use_upsample(w, h, pa - 90, j.filt_type);
      if (need_above && up_a) upsample_wave(above, tmp, w + (need_right ? h : 0), maxv, lane, sync);
      up_l = use_upsample(h, w, pa - 180, j.filt_type);
      if (need_left && up_l) upsample_wave(left, tmp, h + (need_bottom ? w : 0), maxv, lane, sync);
    }
    dr_predict_edges(pa, above, left, w, h, up_a ? 1 : 0, up_l ? 1 : 0, maxv, out, lane);
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
  sync();

  if (j.fi_mode < 5) {
    // 33x33 staging buffer; sub-block (i, k) of 4x2 samples at rows 1+2i, cols 1+4k.
    for (int i = lane; i < (h + 1) * 33; i += 64) fbuf[i] = 0;
    sync();
    if (lane == 0) fbuf[0] = (int16_t)tl;
    for (int c = lane; c < w; c += 64) fbuf[c + 1] = (int16_t)above[c];
    for (int r = lane; r < h; r += 64) fbuf[(r + 1) * 33] = (int16_t)left[r];
    sync();
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
        fbuf[(r + (k >> 2)) * 33 + c + (k & 3)] = (int16_t)clampi(rv, 0, maxv);
      }
      sync();
    }
    for (int i = lane; i < n; i += 64) {
      const int r = i / w, c = i - r * w;
      out(i, (int32_t)fbuf[(r + 1) * 33 + c + 1]);
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
    out(i, v);
  }
}

// Reference samples straight from the plane.
template <typename Pix>
struct PlaneNb {
  const Pix* plane;
  int stride;
  __device__ __forceinline__ int32_t operator()(int x, int y) const { return (int32_t)plane[(long)y * stride + x]; }
};
struct GlobalOut {
  uint16_t* dst;
  __device__ __forceinline__ void operator()(int i, int32_t v) const { dst[i] = (uint16_t)v; }
};

// Jobs of intra_pred_batch: a table of IntraJob records.
struct JobTable {
  const IntraJob* jobs;
  __device__ __forceinline__ IntraJob operator()(int i) const { return jobs[i]; }
  __device__ __forceinline__ bool valid(const IntraJob&) const { return true; }
};

template <typename Pix, class JOBS>
__global__ __launch_bounds__(64) void k_intra_pred(const Pix* __restrict__ plane, int stride, int bd, JOBS jobs,
                                                   int n_jobs, uint16_t* __restrict__ out) {
  __shared__ int16_t above[kEdge], left[kEdge], tmp[kEdge];
  __shared__ int16_t fbuf[33 * 33];
  const int ji = blockIdx.x;
  if (ji >= n_jobs) return;  // whole workgroup exits together
  const IntraJob j = jobs(ji);
  if (!jobs.valid(j)) {  // a survivor of a unit whose reference reads would leave the plane
    for (int i = threadIdx.x; i < j.w * j.h; i += 64) out[j.out_off + i] = 0;
    return;
  }
  predict_wave(j, bd, PlaneNb<Pix>{plane, stride}, GlobalOut{out + j.out_off}, BlockSync{}, above, left, tmp, fbuf,
               (int)threadIdx.x);
}

// ---------------------------------------------------------------------------
// Fused MDS0 fast loop: one workgroup (4 waves) per unit.  The unit's source
// block and reference samples are loaded into LDS once.  Each wave then works
// alone on batches of candidates: it predicts a batch into its private LDS
// residual buffer (source - prediction, 16-bit, tile-major) and runs the
// Hadamard SATD of the batch with 8 lanes per 8x8 sub-tile (a column pass
// and a row pass through LDS), adding each tile's SATD into its candidate's
// LDS counter.  Only wave-level barriers after the loads; predictions never
// leave LDS.  Bit-exact with intra_pred_batch followed by pred_satd_batch.
// ---------------------------------------------------------------------------

// Reference samples of one unit cached in LDS: row[0] = top-left, row[1 + i]
// = above sample i, col[i] = left sample i.
struct CacheNb {
  const int32_t* row;
  const int32_t* col;
  int x0, y0;
  __device__ __forceinline__ int32_t operator()(int x, int y) const {
    return y == y0 - 1 ? row[x - x0 + 1] : col[y - y0];
  }
};

// Residual store of one candidate: sample i of the W x H raster goes to its
// T x T tile (tile-major, raster inside the tile).
template <int W, int H, int T>
struct TileResOut {
  int16_t* res;
  const int16_t* src;
  __device__ __forceinline__ void operator()(int i, int32_t v) const {
    const int r = i / W, c = i % W;
    const int k = (r / T) * (W / T) + c / T;
    res[k * (T * T) + (r % T) * T + (c % T)] = (int16_t)(src[i] - (int16_t)v);
  }
};

// Geometry of the fused kernel for one unit size.
template <int W, int H>
struct MdGeom {
  static constexpr int N = W * H;
  static constexpr int T = (W < H ? W : H) < 32 ? (W < H ? W : H) : 32;  // Hadamard tile side
  static constexpr int TPJ = N / (T * T);                                 // tiles per candidate
  // lanes one candidate's SATD occupies: 1 per 4x4 tile, 8 per 8x8 sub-tile
  static constexpr int LANES = T == 4 ? TPJ : TPJ * (T / 8) * (T / 8) * 8;
  static constexpr int BATCH = LANES >= 64 ? 1 : 64 / LANES;  // candidates per wave batch
  static constexpr int RES = BATCH * N;                       // int16 residual samples per wave
  static constexpr bool FI = W <= 32 && H <= 32;              // filter intra possible
};

// SATD of nt T x T residual tiles at res (contiguous, tile-major) by one
// wave; tile i adds into acc[i / tpj].  The tiles are overwritten.
template <int T>
__device__ void wave_satd_tiles(int16_t* res, int nt, int tpj, uint32_t* acc, int lane) {
  const WaveSync wsync{};
  if constexpr (T == 4) {
    for (int i = lane; i < nt; i += 64) {
      int32_t c[16];
      had4_tile(PackedRes{res + i * 16, 4}, c);
      int32_t sum = 0;
#pragma unroll
      for (int k = 0; k < 16; k++) sum += c[k] < 0 ? -c[k] : c[k];
      atomicAdd(acc + i / tpj, (uint32_t)sum);
    }
  } else {
    constexpr int SUB = (T / 8) * (T / 8);        // 8x8 sub-tiles per tile
    constexpr int LT = SUB * 8;                   // lane slots per tile
    constexpr int TPR = LT >= 64 ? 1 : 64 / LT;   // tiles per round
    constexpr int J = LT > 64 ? LT / 64 : 1;      // slots per lane
    for (int t0 = 0; t0 < nt; t0 += TPR) {
      int16_t* base[J];
      int q[J], L[J], ti[J];
      bool ok[J];
      int16_t o[J][8];
#pragma unroll
      for (int jj = 0; jj < J; jj++) {
        const int sl = lane + 64 * jj;
        ti[jj] = t0 + sl / LT;
        ok[jj] = ti[jj] < nt;
        q[jj] = (sl % LT) >> 3;
        L[jj] = sl & 7;
        base[jj] = res + (ok[jj] ? ti[jj] : 0) * (T * T);
      }
      // column pass: slot (q, L) transforms column L of sub-tile q
#pragma unroll
      for (int jj = 0; jj < J; jj++) {
        if (!ok[jj]) continue;
        int r0, c0;
        if constexpr (T == 8) {
          r0 = 0; c0 = 0;
        } else if constexpr (T == 16) {
          r0 = (q[jj] >> 1) * 8; c0 = (q[jj] & 1) * 8;
        } else {
          const int Q = q[jj] >> 2, qq = q[jj] & 3;
          r0 = (Q >> 1) * 16 + (qq >> 1) * 8; c0 = (Q & 1) * 16 + (qq & 1) * 8;
        }
        int32_t col[8];
#pragma unroll
        for (int r = 0; r < 8; r++) col[r] = base[jj][(r0 + r) * T + c0 + L[jj]];
        had8_col(col, o[jj]);
      }
      wsync();
#pragma unroll
      for (int jj = 0; jj < J; jj++)
        if (ok[jj]) {
#pragma unroll
          for (int k = 0; k < 8; k++) base[jj][q[jj] * 64 + L[jj] * 8 + k] = o[jj][k];
        }
      wsync();
      // row pass: slot (q, L) produces coefficients L*8 .. L*8+7 of sub-tile q
      int32_t sum = 0;
#pragma unroll
      for (int jj = 0; jj < J; jj++) {
        if (!ok[jj]) continue;
        int32_t col[8];
#pragma unroll
        for (int k = 0; k < 8; k++) col[k] = base[jj][q[jj] * 64 + L[jj] + 8 * k];
        had8_col(col, o[jj]);
        if constexpr (T == 8) {
#pragma unroll
          for (int k = 0; k < 8; k++) sum += abs((int32_t)o[jj][k]);
        }
      }
      if constexpr (T >= 16) {
        wsync();
#pragma unroll
        for (int jj = 0; jj < J; jj++)
          if (ok[jj]) {
#pragma unroll
            for (int k = 0; k < 8; k++) base[jj][q[jj] * 64 + L[jj] * 8 + k] = o[jj][k];
          }
        wsync();
        // 16x16 combine (int16 wrap, >> 1): each slot takes 2 (group, position) pairs
#pragma unroll
        for (int jj = 0; jj < J; jj++) {
          if (!ok[jj]) continue;
          const int sl = (lane + 64 * jj) % LT;
#pragma unroll
          for (int e = 0; e < 2; e++) {
            const int pidx = sl * 2 + e;  // 0 .. 64 * SUB / 4 - 1
            int16_t* B = base[jj] + (pidx >> 6) * 256;
            const int ii = pidx & 63;
            const int16_t a0 = B[ii], a1 = B[ii + 64], a2 = B[ii + 128], a3 = B[ii + 192];
            const int16_t b0 = (int16_t)((int16_t)(a0 + a1) >> 1), b1 = (int16_t)((int16_t)(a0 - a1) >> 1);
            const int16_t b2 = (int16_t)((int16_t)(a2 + a3) >> 1), b3 = (int16_t)((int16_t)(a2 - a3) >> 1);
            const int16_t o0 = (int16_t)(b0 + b2), o1 = (int16_t)(b1 + b3);
            const int16_t o2 = (int16_t)(b0 - b2), o3 = (int16_t)(b1 - b3);
            if constexpr (T == 16) {
              sum += abs((int32_t)o0) + abs((int32_t)o1) + abs((int32_t)o2) + abs((int32_t)o3);
            } else {
              B[ii] = o0; B[ii + 64] = o1; B[ii + 128] = o2; B[ii + 192] = o3;
            }
          }
        }
        if constexpr (T == 32) {
          wsync();
          // 32x32 combine (saturating >> 2): each slot takes 2 of the 256 positions
#pragma unroll
          for (int jj = 0; jj < J; jj++) {
            if (!ok[jj]) continue;
            const int sl = (lane + 64 * jj) % LT;
#pragma unroll
            for (int e = 0; e < 2; e++) {
              const int ii = sl * 2 + e;
              const int16_t* Lb = base[jj];
              const int32_t a0 = Lb[ii], a1 = Lb[ii + 256], a2 = Lb[ii + 512], a3 = Lb[ii + 768];
              const int16_t b0 = sat16((a0 + a1) >> 2), b1 = sat16((a0 - a1) >> 2);
              const int16_t b2 = sat16((a2 + a3) >> 2), b3 = sat16((a2 - a3) >> 2);
              sum += abs((int32_t)(int16_t)(b0 + b2)) + abs((int32_t)(int16_t)(b1 + b3)) +
                     abs((int32_t)(int16_t)(b0 - b2)) + abs((int32_t)(int16_t)(b1 - b3));
            }
          }
        }
      }
      // every slot of a tile belongs to tile ti[0] (J > 1 only when a tile spans the wave)
      if (ok[0] && sum != 0) atomicAdd(acc + ti[0] / tpj, (uint32_t)sum);
      wsync();
    }
  }
}

// A unit whose reference reads would leave the plane (availability counts
// inconsistent with its position) gets kSatdInvalid for every candidate.
__device__ __forceinline__ bool md_unit_ok(const MdUnit& u, int w, int h, int pw, int ph, int* na, int* nl) {
  const int nt = u.n_top_px, ntr = u.n_topright_px, nlf = u.n_left_px, nbl = u.n_bottomleft_px;
  if (nt < 0 || ntr < 0 || nlf < 0 || nbl < 0) return false;
  *na = nt > 0 ? max(nt, ntr > 0 ? w + ntr : 0) : 0;
  *nl = nlf > 0 ? max(nlf, nbl > 0 ? h + nbl : 0) : 0;
  return u.x >= 0 && u.y >= 0 && u.x <= pw - w && u.y <= ph - h && (nt == 0 || u.y >= 1) && (nlf == 0 || u.x >= 1) &&
         u.x + *na <= pw && u.y + *nl <= ph && *na <= kEdge - 1 && *nl <= kEdge;
}

// Shared directional edges of a unit.  A directional candidate's final
// edges depend on its angle only through its zone (which edges and
// extensions it needs) and its corner-filter / edge-strength / upsample
// choices, so the unit's candidates fall into a few edge configurations.
// The unit's raw edges are assembled once (with the top-right and
// bottom-left extensions: a candidate that needs no extension never reads
// past the unextended length, and the edge filters stop at their own
// counts), then every configuration's filtered / upsampled edges once; the
// candidates then only predict.  Key: zone (1..3) | corner << 2 |
// above strength << 3 | left strength << 5 | upsample above << 7 |
// upsample left << 8.
constexpr int kMdSlots = 16;
constexpr uint16_t kKeyNonDir = 0xffff, kKeyConst = 0xfffe;

template <int W, int H>
__device__ __forceinline__ uint16_t md_edge_key(const MdCand& cd, const MdUnit& U, int fi_ok) {
  const bool dir = (cd.mode >= 3 && cd.mode <= 8) || ((cd.mode == 1 || cd.mode == 2) && cd.delta != 0);
  if (!dir || (fi_ok && cd.fi_mode < 5)) return kKeyNonDir;
  // an extension without the full edge beside it (not produced by AV1's
  // availability rules) makes the extended fill differ by zone: own edges
  if ((U.n_topright_px > 0 && U.n_top_px < W) || (U.n_bottomleft_px > 0 && U.n_left_px < H)) return kKeyNonDir;
  const int pa = MODE_TO_ANGLE[cd.mode] + 3 * cd.delta;
  const bool na = pa < 180, nl = pa > 90;
  if ((!na && U.n_left_px == 0) || (!nl && U.n_top_px == 0)) return kKeyConst;
  const int zone = pa < 90 ? 1 : (pa < 180 ? 2 : 3);
  int cf = 0, sa = 0, sl = 0, ua = 0, ul = 0;
  if (U.edge_filter) {
    cf = na && nl && (W + H >= 24);
    if (na && U.n_top_px > 0) sa = edge_strength(W, H, pa - 90, U.filt_type);
    if (nl && U.n_left_px > 0) sl = edge_strength(H, W, pa - 180, U.filt_type);
    ua = na && use_upsample(W, H, pa - 90, U.filt_type);
    ul = nl && use_upsample(H, W, pa - 180, U.filt_type);
  }
  return (uint16_t)(zone | cf << 2 | sa << 3 | sl << 5 | ua << 7 | ul << 8);
}

// Filtered / upsampled edges of configuration `key` from the raw edges, by
// one wave (tmp: the wave's scratch).
template <int W, int H>
__device__ void md_build_slot(uint16_t key, const int16_t* rawA, const int16_t* rawL, int16_t* A, int16_t* Lf,
                              int16_t* tmp, const MdUnit& U, int maxv, int lane) {
  const WaveSync sync{};
  const int zone = key & 3;
  const bool na = zone <= 2, nl = zone >= 2, nr = zone == 1, nbm = zone == 3;
  const int cf = (key >> 2) & 1, sa = (key >> 3) & 3, sl = (key >> 5) & 3, ua = (key >> 7) & 1, ul = (key >> 8) & 1;
  for (int i = lane; i < kEdge; i += 64) {
    A[i] = rawA[i];
    Lf[i] = rawL[i];
  }
  sync();
  if (cf) {
    if (lane == 0) {
      const int32_t v = (Lf[kO] * 5 + A[kO - 1] * 6 + A[kO] * 5 + 8) >> 4;
      A[kO - 1] = (int16_t)v;
      Lf[kO - 1] = (int16_t)v;
    }
    sync();
  }
  if (na && sa) filter_edge_wave(A, tmp, kO - 1, U.n_top_px + 1 + (nr ? H : 0), sa, lane, sync);
  if (nl && sl) filter_edge_wave(Lf, tmp, kO - 1, U.n_left_px + 1 + (nbm ? W : 0), sl, lane, sync);
  if (na && ua) upsample_wave(A, tmp, W + (nr ? H : 0), maxv, lane, sync);
  if (nl && ul) upsample_wave(Lf, tmp, H + (nbm ? W : 0), maxv, lane, sync);
}

template <typename Pix, int W, int H>
__global__ __launch_bounds__(256) void k_md_fast_satd(const Pix* __restrict__ src, const Pix* __restrict__ recon,
                                                      int stride, int pw, int ph, int bd,
                                                      const MdUnit* __restrict__ units, int n_units, MdCandList cl,
                                                      uint32_t* __restrict__ satd_out) {
  using G = MdGeom<W, H>;
  __shared__ int16_t res[4][G::RES];
  __shared__ int16_t srcb[G::N];
  __shared__ int32_t rowc[kEdge], colc[kEdge];
  __shared__ int16_t above[4][kEdge], left[4][kEdge], tmp[4][kEdge];
  __shared__ int16_t fbuf[4][G::FI ? (H + 1) * 33 : 1];
  __shared__ uint32_t acc[64];
  __shared__ int16_t rawA[kEdge], rawL[kEdge];
  __shared__ int16_t slotA[kMdSlots][kEdge], slotL[kMdSlots][kEdge];
  __shared__ uint16_t skey[kMdSlots];
  __shared__ uint8_t cslot[64];  // candidate -> slot; 0xff: own edges, 0xfe: constant fill
  __shared__ int nslot_sh;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int u = xcd_chunked(blockIdx.x, gridDim.x);
  if (u >= n_units) return;
  const MdUnit U = units[u];
  const int nc = cl.n;
  int na = 0, nl = 0;
  if (!md_unit_ok(U, W, H, pw, ph, &na, &nl)) {
    for (int m = tid; m < nc; m += 256) satd_out[(long)u * nc + m] = kSatdInvalid;
    return;
  }
  for (int i = tid; i < G::N; i += 256) srcb[i] = (int16_t)src[(long)(U.y + i / W) * stride + U.x + i % W];
  if (tid == 0) rowc[0] = (U.n_top_px > 0 && U.n_left_px > 0) ? (int32_t)recon[(long)(U.y - 1) * stride + U.x - 1] : 0;
  for (int i = tid; i < na; i += 256) rowc[1 + i] = recon[(long)(U.y - 1) * stride + U.x + i];
  for (int i = tid; i < nl; i += 256) colc[i] = recon[(long)(U.y + i) * stride + U.x - 1];
  for (int m = tid; m < nc; m += 256) acc[m] = 0;
  __syncthreads();

  const CacheNb nb{rowc, colc, U.x, U.y};
  const int base = 128 << (bd - 8), maxv = (1 << bd) - 1;
  // ---- the unit's raw directional edges (with both extensions) ----
  {
    const int n_top = U.n_top_px, n_tr = U.n_topright_px, n_left = U.n_left_px, n_bl = U.n_bottomleft_px;
    for (int i = tid; i < kEdge; i += 256) {
      int32_t a = base - 1, l = base + 1;
      const int k = i - kO;
      if (k >= 0 && k < W + H) {
        if (n_top > 0) a = nb(U.x + min(k, (n_tr > 0 ? W + n_tr : n_top) - 1), U.y - 1);
        else if (n_left > 0) a = nb(U.x - 1, U.y);
        if (n_left > 0) l = nb(U.x - 1, U.y + min(k, (n_bl > 0 ? H + n_bl : n_left) - 1));
        else if (n_top > 0) l = nb(U.x, U.y - 1);
      } else if (k == -1) {
        const int32_t tl = (n_top > 0 && n_left > 0) ? nb(U.x - 1, U.y - 1)
                         : n_top > 0 ? nb(U.x, U.y - 1) : (n_left > 0 ? nb(U.x - 1, U.y) : base);
        a = tl;
        l = tl;
      }
      rawA[i] = (int16_t)a;
      rawL[i] = (int16_t)l;
    }
  }
  // ---- edge configurations: candidate keys, distinct keys -> slots ----
  if (wave == 0) {
    const uint16_t key = lane < nc ? md_edge_key<W, H>(cl.c[lane], U, G::FI) : kKeyNonDir;
    uint64_t todo = __ballot(key < kKeyConst);
    int ns = 0;
    uint8_t my = key == kKeyConst ? 0xfe : 0xff;
    while (todo) {
      const int lead = __ffsll((unsigned long long)todo) - 1;
      const uint16_t k = (uint16_t)__shfl((int)key, lead, 64);
      const bool mine = key == k;
      if (ns < kMdSlots) {
        if (mine) my = (uint8_t)ns;
        if (lane == 0) skey[ns] = k;
        ns++;
      }
      todo &= ~__ballot(mine);
    }
    if (lane < nc) cslot[lane] = my;
    if (lane == 0) nslot_sh = ns;
  }
  __syncthreads();
  const int nslot = nslot_sh;
  for (int sl = wave; sl < nslot; sl += 4)
    md_build_slot<W, H>(skey[sl], rawA, rawL, slotA[sl], slotL[sl], tmp[wave], U, maxv, lane);
  __syncthreads();
  // batches of at most BATCH candidates, spread so every wave gets work
  const int per = (nc + 3) / 4;
  const int bsz = per < G::BATCH ? per : G::BATCH;
  for (int m0 = wave * bsz; m0 < nc; m0 += 4 * bsz) {
    const int nb_c = min(bsz, nc - m0);
    for (int b = 0; b < nb_c; b++) {
      const MdCand cd = cl.c[m0 + b];
      const TileResOut<W, H, G::T> ro{res[wave] + b * G::N, srcb};
      const int cs = cslot[m0 + b];
      if (cs < kMdSlots) {  // directional, shared edges
        const uint16_t key = skey[cs];
        dr_predict_edges(MODE_TO_ANGLE[cd.mode] + 3 * cd.delta, slotA[cs], slotL[cs], W, H, (key >> 7) & 1,
                         (key >> 8) & 1, maxv, ro, lane);
        continue;
      }
      if (cs == 0xfe) {  // directional without the edge it needs: one value
        const int pa = MODE_TO_ANGLE[cd.mode] + 3 * cd.delta;
        const int32_t v = pa > 90 ? (U.n_top_px > 0 ? nb(U.x, U.y - 1) : base + 1)
                                  : (U.n_left_px > 0 ? nb(U.x - 1, U.y) : base - 1);
        for (int i = lane; i < G::N; i += 64) ro(i, v);
        continue;
      }
      IntraJob j;
      j.x = U.x; j.y = U.y; j.w = (uint8_t)W; j.h = (uint8_t)H;
      j.mode = cd.mode; j.delta = cd.delta; j.fi_mode = G::FI ? cd.fi_mode : (uint8_t)5;
      j.edge_filter = U.edge_filter; j.filt_type = U.filt_type; j.pad0 = 0;
      j.n_top_px = U.n_top_px; j.n_topright_px = U.n_topright_px;
      j.n_left_px = U.n_left_px; j.n_bottomleft_px = U.n_bottomleft_px;
      j.out_off = 0; j.pad1 = 0;
      predict_wave<W, H>(j, bd, nb, ro, WaveSync{}, above[wave], left[wave], tmp[wave], fbuf[wave], lane);
      WaveSync{}();
    }
    WaveSync{}();
    wave_satd_tiles<G::T>(res[wave], nb_c * G::TPJ, G::TPJ, acc + m0, lane);
  }
  __syncthreads();
  for (int m = tid; m < nc; m += 256) satd_out[(long)u * nc + m] = acc[m];
}

// Survivor predictions: job i = (unit i / k, rank i % k) predicts candidate
// sel[i] of its unit into out + i * w * h.
struct SurvivorJobs {
  const MdUnit* units;
  MdCandList cl;
  const uint8_t* sel;
  int k, w, h, pw, ph;
  __device__ __forceinline__ bool valid(const IntraJob& j) const {
    MdUnit u;
    u.x = j.x; u.y = j.y;
    u.n_top_px = j.n_top_px; u.n_topright_px = j.n_topright_px;
    u.n_left_px = j.n_left_px; u.n_bottomleft_px = j.n_bottomleft_px;
    int na, nl;
    return md_unit_ok(u, w, h, pw, ph, &na, &nl);
  }
  __device__ __forceinline__ IntraJob operator()(int i) const {
    const MdUnit U = units[i / k];
    const MdCand cd = cl.c[sel[i] < cl.n ? sel[i] : 0];
    IntraJob j;
    j.x = U.x; j.y = U.y; j.w = (uint8_t)w; j.h = (uint8_t)h;
    j.mode = cd.mode; j.delta = cd.delta; j.fi_mode = cd.fi_mode;
    j.edge_filter = U.edge_filter; j.filt_type = U.filt_type; j.pad0 = 0;
    j.n_top_px = U.n_top_px; j.n_topright_px = U.n_topright_px;
    j.n_left_px = U.n_left_px; j.n_bottomleft_px = U.n_bottomleft_px;
    j.out_off = (uint32_t)i * (uint32_t)(w * h); j.pad1 = 0;
    return j;
  }
};

template <typename Pix>
static hipError_t launch_md_satd(const Pix* src, const Pix* recon, int stride, int pw, int ph, int bd, int w, int h,
                                 const MdUnit* units, int n_units, const MdCandList& cl, uint32_t* satd,
                                 hipStream_t st) {
#define X(W, H)                                                                                                  \
  if (w == W && h == H) {                                                                                        \
    hipLaunchKernelGGL((k_md_fast_satd<Pix, W, H>), dim3(n_units), dim3(256), 0, st, src, recon, stride, pw, ph, \
                       bd, units, n_units, cl, satd);                                                            \
    return hipGetLastError();                                                                                    \
  }
  X(4, 4) X(8, 8) X(16, 16) X(32, 32) X(64, 64) X(4, 8) X(8, 4) X(8, 16) X(16, 8) X(16, 32) X(32, 16) X(32, 64)
  X(64, 32) X(4, 16) X(16, 4) X(8, 32) X(32, 8) X(16, 64) X(64, 16)
#undef X
  return hipErrorInvalidValue;
}

hipError_t md_fast_satd(const void* src, const void* recon, int stride, int pw, int ph, int bd, int w, int h,
                        const MdUnit* units, int n_units, const MdCandList& cl, uint32_t* satd, hipStream_t st) {
  if (n_units <= 0) return hipSuccess;
  if (bd == 8)
    return launch_md_satd((const uint8_t*)src, (const uint8_t*)recon, stride, pw, ph, bd, w, h, units, n_units, cl,
                          satd, st);
  return launch_md_satd((const uint16_t*)src, (const uint16_t*)recon, stride, pw, ph, bd, w, h, units, n_units, cl,
                        satd, st);
}

hipError_t md_predict_survivors(const void* recon, int stride, int pw, int ph, int bd, int w, int h,
                                const MdUnit* units, int n_units, const MdCandList& cl, const uint8_t* sel, int k,
                                uint16_t* out, hipStream_t st) {
  const long n = (long)n_units * k;
  if (n <= 0) return hipSuccess;
  const SurvivorJobs sj{units, cl, sel, k, w, h, pw, ph};
  if (bd == 8)
    hipLaunchKernelGGL((k_intra_pred<uint8_t, SurvivorJobs>), dim3((unsigned)n), dim3(64), 0, st,
                       (const uint8_t*)recon, stride, bd, sj, (int)n, out);
  else
    hipLaunchKernelGGL((k_intra_pred<uint16_t, SurvivorJobs>), dim3((unsigned)n), dim3(64), 0, st,
                       (const uint16_t*)recon, stride, bd, sj, (int)n, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain form: one unit predicted from edges the caller supplies.
// kind 0 (predict_dc / _v / _h / _smooth* / _paeth, predict_filter_intra):
// the predictor reads its reference samples from the cached row / column;
// kind 1 (dr_predictor_edged): the final edge buffers, origin kO.
// ---------------------------------------------------------------------------
template <typename Pix>
struct PackedOut {
  Pix* dst;
  __device__ __forceinline__ void operator()(int i, int32_t v) const { dst[i] = (Pix)v; }
};

template <typename Pix>
__global__ __launch_bounds__(64) void k_plain_intra(PlainIntra p, const int32_t* __restrict__ row,
                                                    const int32_t* __restrict__ col,
                                                    const int16_t* __restrict__ above_e,
                                                    const int16_t* __restrict__ left_e, Pix* __restrict__ out) {
  __shared__ int16_t above[kEdge], left[kEdge], tmp[kEdge];
  __shared__ int16_t fbuf[33 * 33];
  const int lane = threadIdx.x;
  const int maxv = (1 << p.bd) - 1;
  if (p.kind == 1) {
    for (int i = lane; i < kEdge; i += 64) {
      above[i] = above_e[i];
      left[i] = left_e[i];
    }
    __syncthreads();
    dr_predict_edges(p.angle, above, left, p.w, p.h, p.ua, p.ul, maxv, PackedOut<Pix>{out}, lane);
    return;
  }
  IntraJob j;
  j.x = 1; j.y = 1; j.w = (uint8_t)p.w; j.h = (uint8_t)p.h;
  j.mode = (uint8_t)p.mode; j.delta = 0; j.fi_mode = (uint8_t)p.fi_mode;
  j.edge_filter = 0; j.filt_type = 0; j.pad0 = 0;
  j.n_top_px = (int16_t)(p.has_above ? p.w : 0); j.n_topright_px = 0;
  j.n_left_px = (int16_t)(p.has_left ? p.h : 0); j.n_bottomleft_px = 0;
  j.out_off = 0; j.pad1 = 0;
  predict_wave(j, p.bd, CacheNb{row, col, 1, 1}, PackedOut<Pix>{out}, BlockSync{}, above, left, tmp, fbuf, lane);
}

template <typename Pix>
__global__ __launch_bounds__(256) void k_plain_cfl(const int16_t* __restrict__ ac, const Pix* __restrict__ pred,
                                                   int alpha_q3, int n, int maxv, Pix* __restrict__ out) {
  for (int i = threadIdx.x; i < n; i += 256) {
    const int32_t sc = alpha_q3 * ac[i];
    const int32_t q0 = sc < 0 ? -((-sc + 32) >> 6) : ((sc + 32) >> 6);
    out[i] = (Pix)clampi(q0 + pred[i], 0, maxv);
  }
}

hipError_t plain_intra(const PlainIntra& p, const int32_t* row, const int32_t* col, const int16_t* above,
                       const int16_t* left, void* out, hipStream_t st) {
  if (p.bd > 8)
    hipLaunchKernelGGL((k_plain_intra<uint16_t>), dim3(1), dim3(64), 0, st, p, row, col, above, left,
                       (uint16_t*)out);
  else
    hipLaunchKernelGGL((k_plain_intra<uint8_t>), dim3(1), dim3(64), 0, st, p, row, col, above, left, (uint8_t*)out);
  return hipGetLastError();
}

hipError_t plain_cfl(const int16_t* ac, const void* pred, int alpha_q3, int n, int bd, void* out, hipStream_t st) {
  const int maxv = (1 << bd) - 1;
  if (bd > 8)
    hipLaunchKernelGGL((k_plain_cfl<uint16_t>), dim3(1), dim3(256), 0, st, ac, (const uint16_t*)pred, alpha_q3, n,
                       maxv, (uint16_t*)out);
  else
    hipLaunchKernelGGL((k_plain_cfl<uint8_t>), dim3(1), dim3(256), 0, st, ac, (const uint8_t*)pred, alpha_q3, n,
                       maxv, (uint8_t*)out);
  return hipGetLastError();
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
    hipLaunchKernelGGL((k_intra_pred<uint8_t, JobTable>), dim3(n_jobs), dim3(64), 0, st, (const uint8_t*)plane,
                       stride, bd, JobTable{jobs}, n_jobs, out);
  else
    hipLaunchKernelGGL((k_intra_pred<uint16_t, JobTable>), dim3(n_jobs), dim3(64), 0, st, (const uint16_t*)plane,
                       stride, bd, JobTable{jobs}, n_jobs, out);
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
