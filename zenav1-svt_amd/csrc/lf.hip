// Loop-filter stage kernels for gfx950: deblocking, CDEF direction +
// strength-SSE search, Wiener covariance and self-guided projection
// statistics per restoration unit, and the per-unit SSE.
//
// Semantics:
//   deblock   rust/crates/svtav1-dsp/src/loop_filter.rs:50-405, hbd.rs:1040-1433,
//             rust/crates/svtav1-encoder/src/deblock.rs:698 lpf_params, :811 filter_plane.
//             All vertical edges of a plane, then all horizontal edges.  An
//             edge's filter reaches at most half of the smaller of the two
//             transforms it separates, so no two edges of one pass touch the
//             same sample and each pass runs fully in parallel.
//   CDEF      rust/crates/svtav1-dsp/src/cdef.rs:113-1228,
//             rust/crates/svtav1-encoder/src/cdef.rs:517,952,1045,1128,1166,1374
//   Wiener    rust/crates/svtav1-dsp/src/restoration.rs:198,232,567,1388,1395,1707,1738
//   unit SSE  rust/crates/svtav1-dsp/src/restoration.rs:1666,1991
//   SGR       self-guided filter + projection normal equations (no reference
//             twin; AV1 self-guided filter process, fast r = 2 variant)
#include <algorithm>
#include <type_traits>

#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {

namespace {
__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// ---------------------------------------------------------------------------
// deblocking
// ---------------------------------------------------------------------------
struct Thr {
  int32_t lim, mblim, hev, shift, bd;
};
struct LevelList {
  uint8_t l[64];
};

__device__ __forceinline__ int32_t sclamp(int32_t t, int bd) {
  const int32_t m = 128 << (bd - 8);
  return clampi(t, -m, m - 1);
}

__device__ __forceinline__ void filter4(bool mask, const Thr& t, int32_t* w) {
  const int32_t bias = 0x80 << t.shift;
  const int32_t ps1 = w[0] - bias, ps0 = w[1] - bias, qs0 = w[2] - bias, qs1 = w[3] - bias;
  const int32_t th = t.hev << t.shift;
  const int32_t hev = (abs(w[0] - w[1]) > th || abs(w[3] - w[2]) > th) ? -1 : 0;
  int32_t f = sclamp(ps1 - qs1, t.bd) & hev;
  f = sclamp(f + 3 * (qs0 - ps0), t.bd) & (mask ? -1 : 0);
  const int32_t f1 = sclamp(f + 4, t.bd) >> 3, f2 = sclamp(f + 3, t.bd) >> 3;
  w[2] = sclamp(qs0 - f1, t.bd) + bias;
  w[1] = sclamp(ps0 + f2, t.bd) + bias;
  const int32_t o = ((f1 + 1) >> 1) & ~hev;
  w[3] = sclamp(qs1 - o, t.bd) + bias;
  w[0] = sclamp(ps1 + o, t.bd) + bias;
}

// Filter mask over N samples (N = 4, 6, 8) centred on the edge.
template <int N>
__device__ __forceinline__ bool mask_n(const Thr& t, const int32_t* w) {
  const int32_t l = t.lim << t.shift, bl = t.mblim << t.shift;
  bool bad = false;
#pragma unroll
  for (int k = 0; k + 1 < N / 2; k++) bad |= abs(w[k] - w[k + 1]) > l;
#pragma unroll
  for (int k = N / 2; k + 1 < N; k++) bad |= abs(w[k] - w[k + 1]) > l;
  bad |= abs(w[N / 2 - 1] - w[N / 2]) * 2 + abs(w[N / 2 - 2] - w[N / 2 + 1]) / 2 > bl;
  return !bad;
}

// Flatness with p0 = w[c-1], q0 = w[c]: taps k = K0 .. K1-1 on each side.
template <int K0, int K1>
__device__ __forceinline__ bool flat_at(int shift, const int32_t* w, int c) {
  const int32_t t = 1 << shift;
  bool bad = false;
#pragma unroll
  for (int k = K0; k < K1; k++) bad |= abs(w[c - 1 - k] - w[c - 1]) > t || abs(w[c + k] - w[c]) > t;
  return !bad;
}

#define R3(x) (((x) + 4) >> 3)
#define R4(x) (((x) + 8) >> 4)

template <int LEN>
__device__ __forceinline__ void lpf_line(const Thr& t, int32_t* w) {
  if constexpr (LEN == 4) {
    filter4(mask_n<4>(t, w), t, w);
  } else if constexpr (LEN == 6) {
    const bool m = mask_n<6>(t, w);
    if (m && flat_at<1, 3>(t.shift, w, 3)) {
      const int32_t p2 = w[0], p1 = w[1], p0 = w[2], q0 = w[3], q1 = w[4], q2 = w[5];
      w[1] = R3(p2 * 3 + p1 * 2 + p0 * 2 + q0);
      w[2] = R3(p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1);
      w[3] = R3(p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2);
      w[4] = R3(p0 + q0 * 2 + q1 * 2 + q2 * 3);
    } else {
      filter4(m, t, w + 1);
    }
  } else if constexpr (LEN == 8) {
    const bool m = mask_n<8>(t, w);
    if (m && flat_at<1, 4>(t.shift, w, 4)) {
      const int32_t p3 = w[0], p2 = w[1], p1 = w[2], p0 = w[3], q0 = w[4], q1 = w[5], q2 = w[6], q3 = w[7];
      w[1] = R3(p3 + p3 + p3 + 2 * p2 + p1 + p0 + q0);
      w[2] = R3(p3 + p3 + p2 + 2 * p1 + p0 + q0 + q1);
      w[3] = R3(p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2);
      w[4] = R3(p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3);
      w[5] = R3(p1 + p0 + q0 + 2 * q1 + q2 + q3 + q3);
      w[6] = R3(p0 + q0 + q1 + 2 * q2 + q3 + q3 + q3);
    } else {
      filter4(m, t, w + 2);
    }
  } else {
    const bool m = mask_n<8>(t, w + 3);
    const bool fl = flat_at<1, 4>(t.shift, w, 7);
    const bool fl2 = flat_at<4, 7>(t.shift, w, 7);
    if (fl2 && fl && m) {
      const int32_t p6 = w[0], p5 = w[1], p4 = w[2], p3 = w[3], p2 = w[4], p1 = w[5], p0 = w[6];
      const int32_t q0 = w[7], q1 = w[8], q2 = w[9], q3 = w[10], q4 = w[11], q5 = w[12], q6 = w[13];
      w[1] = R4(p6 * 7 + p5 * 2 + p4 * 2 + p3 + p2 + p1 + p0 + q0);
      w[2] = R4(p6 * 5 + p5 * 2 + p4 * 2 + p3 * 2 + p2 + p1 + p0 + q0 + q1);
      w[3] = R4(p6 * 4 + p5 + p4 * 2 + p3 * 2 + p2 * 2 + p1 + p0 + q0 + q1 + q2);
      w[4] = R4(p6 * 3 + p5 + p4 + p3 * 2 + p2 * 2 + p1 * 2 + p0 + q0 + q1 + q2 + q3);
      w[5] = R4(p6 * 2 + p5 + p4 + p3 + p2 * 2 + p1 * 2 + p0 * 2 + q0 + q1 + q2 + q3 + q4);
      w[6] = R4(p6 + p5 + p4 + p3 + p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + q2 + q3 + q4 + q5);
      w[7] = R4(p5 + p4 + p3 + p2 + p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + q3 + q4 + q5 + q6);
      w[8] = R4(p4 + p3 + p2 + p1 + p0 + q0 * 2 + q1 * 2 + q2 * 2 + q3 + q4 + q5 + q6 * 2);
      w[9] = R4(p3 + p2 + p1 + p0 + q0 + q1 * 2 + q2 * 2 + q3 * 2 + q4 + q5 + q6 * 3);
      w[10] = R4(p2 + p1 + p0 + q0 + q1 + q2 * 2 + q3 * 2 + q4 * 2 + q5 + q6 * 4);
      w[11] = R4(p1 + p0 + q0 + q1 + q2 + q3 * 2 + q4 * 2 + q5 * 2 + q6 * 5);
      w[12] = R4(p0 + q0 + q1 + q2 + q3 + q4 * 2 + q5 * 2 + q6 * 7);
    } else if (fl && m) {
      const int32_t p3 = w[3], p2 = w[4], p1 = w[5], p0 = w[6], q0 = w[7], q1 = w[8], q2 = w[9], q3 = w[10];
      w[4] = R3(p3 + p3 + p3 + 2 * p2 + p1 + p0 + q0);
      w[5] = R3(p3 + p3 + p2 + 2 * p1 + p0 + q0 + q1);
      w[6] = R3(p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2);
      w[7] = R3(p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3);
      w[8] = R3(p1 + p0 + q0 + 2 * q1 + q2 + q3 + q3);
      w[9] = R3(p0 + q0 + q1 + 2 * q2 + q3 + q3 + q3);
    } else {
      filter4(m, t, w + 5);
    }
  }
}

// lpf_params (deblock.rs:698): filter length of the edge at (x, y) of a
// plane (x or y a multiple of 4 along the walked axis), 0 = no filtering.
// The reference walks each mi row jumping from transform start to transform
// start; with size-aligned transforms that visits exactly the positions whose
// coordinate is a multiple of the covering transform's size, which is the
// test applied here per position.
__device__ __forceinline__ int edge_len(const DeblockGeomDev& g, bool vert, int x, int y, int plane, int ss,
                                        int level) {
  if (level == 0 || x >= (g.true_w >> ss) || y >= (g.true_h >> ss)) return 0;
  const int mi_row = ss | ((y << ss) >> 2), mi_col = ss | ((x << ss) >> 2);
  const long idx = (long)mi_row * g.mi_cols + mi_col;
  const int ts = vert ? (plane == 0 ? (int)g.tw[idx] : max((int)g.bw[idx] >> 1, 4))
                      : (plane == 0 ? (int)g.th[idx] : max((int)g.bh[idx] >> 1, 4));
  const int coord = vert ? x : y;
  if ((coord & (ts - 1)) || coord == 0) return 0;
  const long prev = vert ? idx - (1 << ss) : idx - ((long)g.mi_cols << ss);
  const int pv = vert ? (plane == 0 ? (int)g.tw[prev] : max((int)g.bw[prev] >> 1, 4))
                      : (plane == 0 ? (int)g.th[prev] : max((int)g.bh[prev] >> 1, 4));
  const int cs = g.skip ? g.skip[idx] : 0, ps = g.skip ? g.skip[prev] : 0;
  const bool pu_edge = g.block_id[prev] != g.block_id[idx];
  if (!(!ps || !cs || pu_edge)) return 0;
  const int md = min(ts, pv);
  return md == 4 ? 4 : (plane != 0 ? 6 : (md == 8 ? 8 : 14));
}

template <typename Pix, int LEN>
__device__ __forceinline__ void filter_at(Pix* buf, long start, long step, const Thr& t) {
  int32_t w[LEN];
#pragma unroll
  for (int k = 0; k < LEN; k++) w[k] = buf[start + k * step];
  lpf_line<LEN>(t, w);
#pragma unroll
  for (int k = 0; k < LEN; k++) buf[start + k * step] = (Pix)w[k];
}

}  // namespace

namespace {
// One (edge position, line) of a deblocking pass: thread `tid` of the pass.
// VERT: consecutive threads take consecutive mi columns of one pixel row;
// HORZ: consecutive pixel columns of one mi row.
template <typename Pix, bool VERT>
__device__ __forceinline__ void deblock_line(Pix* __restrict__ buf, int stride, int pw, int ph, int plane, int ss,
                                             int level, const Thr& t, const DeblockGeomDev& g, long tid) {
  int x, y;
  if (VERT) {
    const int nm = pw >> 2, nl = ph & ~3;
    if (tid >= (long)nm * nl) return;
    y = (int)(tid / nm);
    x = (int)(tid - (long)y * nm) * 4;
  } else {
    const int nm = ph >> 2, nl = pw & ~3;
    if (tid >= (long)nm * nl) return;
    const int my = (int)(tid / nl);
    x = (int)(tid - (long)my * nl);
    y = my * 4;
  }
  const int len = edge_len(g, VERT, VERT ? x : (x & ~3), VERT ? (y & ~3) : y, plane, ss, level);
  if (len == 0) return;
  const long step = VERT ? 1 : stride;
  const long start = (long)y * stride + x - (len / 2) * step;
  if (len == 4) filter_at<Pix, 4>(buf, start, step, t);
  else if (len == 6) filter_at<Pix, 6>(buf, start, step, t);
  else if (len == 8) filter_at<Pix, 8>(buf, start, step, t);
  else filter_at<Pix, 14>(buf, start, step, t);
}

// lf_thresholds (loop_filter.rs:50) at bit depth bd.
__host__ __device__ __forceinline__ Thr make_thr(int level, int sharpness, int bd) {
  int bil = level >> ((sharpness > 0) + (sharpness > 4));
  if (sharpness > 0 && bil > 9 - sharpness) bil = 9 - sharpness;
  if (bil < 1) bil = 1;
  Thr t;
  t.mblim = 2 * (level + 2) + bil;
  t.lim = bil;
  t.hev = level >> 4;
  t.shift = bd - 8;
  t.bd = bd;
  return t;
}
}  // namespace

template <typename Pix, bool VERT>
__global__ __launch_bounds__(256) void k_deblock(Pix* __restrict__ buf, int stride, int pw, int ph, int plane,
                                                 int ss, int level, Thr t, DeblockGeomDev g) {
  deblock_line<Pix, VERT>(buf, stride, pw, ph, plane, ss, level, t, g, (long)blockIdx.x * 256 + threadIdx.x);
}

// Level trials: blockIdx.y selects candidate level levels[y] and its own
// tight plane copy work + y * pw * ph; the level applies to both directions.
template <typename Pix, bool VERT>
__global__ __launch_bounds__(256) void k_deblock_trials(Pix* __restrict__ work, int pw, int ph, int plane, int ss,
                                                        LevelList levels, int sharpness, int bd, DeblockGeomDev g) {
  const int level = levels.l[blockIdx.y];
  if (level == 0) return;
  deblock_line<Pix, VERT>(work + (long)blockIdx.y * pw * ph, pw, pw, ph, plane, ss, level,
                          make_thr(level, sharpness, bd), g, (long)blockIdx.x * 256 + threadIdx.x);
}

template <typename Pix>
__global__ __launch_bounds__(256) void k_trial_copy(const Pix* __restrict__ recon, Pix* __restrict__ work, long n) {
  Pix* dst = work + (long)blockIdx.y * n;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = recon[i];
}

__global__ __launch_bounds__(64) void k_zero_u64(uint64_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// SSE of each trial plane against the source (plane_sse, deblock.rs:208):
// grid-stride partial sums, wave reduction, one atomic per wave.
template <typename Pix>
__global__ __launch_bounds__(256) void k_trial_sse(const Pix* __restrict__ src, const Pix* __restrict__ work, long n,
                                                   unsigned long long* __restrict__ sse) {
  const Pix* a = work + (long)blockIdx.y * n;
  uint64_t s = 0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int32_t d = (int32_t)src[i] - (int32_t)a[i];
    s += (uint64_t)(uint32_t)__mul24(d, d);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&sse[blockIdx.y], (unsigned long long)s);
}

// ---------------------------------------------------------------------------
// CDEF
// ---------------------------------------------------------------------------
constexpr int kVB = 3, kHB = 8;  // vertical / horizontal border of the padded block
constexpr uint16_t kLarge = 0x7f7f;
constexpr int kTS = 80;          // LDS tile stride (64 + 2 * kHB)

// Tap offsets (k = 0, 1) of the 8 directions in the LDS tile, row dir + 2;
// the two extra rows at each end let dir - 2 and dir + 2 index directly.
__constant__ int kCdefDir[12][2] = {
    {kTS, 2 * kTS},  {kTS, 2 * kTS - 1},     {-kTS + 1, -2 * kTS + 2}, {1, -kTS + 2},
    {1, 2},          {1, kTS + 2},           {kTS + 1, 2 * kTS + 2},   {kTS, 2 * kTS + 1},
    {kTS, 2 * kTS},  {kTS, 2 * kTS - 1},     {-kTS + 1, -2 * kTS + 2}, {1, -kTS + 2}};

namespace {
__device__ __forceinline__ int msb32(uint32_t n) { return 31 - __clz(n); }

// constrain(diff, thr, damping) for thr > 0, with sh = max(0, damping - msb(thr)).
__device__ __forceinline__ int32_t constrain(int32_t diff, int32_t thr, int32_t sh) {
  const int32_t ad = abs(diff);
  const int32_t v = min(ad, max(0, thr - (ad >> sh)));
  return diff < 0 ? -v : v;
}

__device__ __forceinline__ int adjust_strength(int s, int var) {
  const int i = (var >> 6) ? min(msb32((uint32_t)(var >> 6)), 12) : 0;
  return var ? (s * (4 + i) + 8) >> 4 : 0;
}

// cdef_find_dir (cdef.rs:143) on the 8x8 at t (row stride ST).
template <int ST = kTS>
__device__ int find_dir(const uint16_t* t, int cs, int* var_out) {
  constexpr int DIV[9] = {0, 840, 420, 280, 210, 168, 140, 120, 105};
  int32_t partial[8][15];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 15; b++) partial[a][b] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int x = (t[i * ST + j] >> cs) - 128;
      partial[0][i + j] += x;
      partial[1][i + j / 2] += x;
      partial[2][i] += x;
      partial[3][3 + i - j / 2] += x;
      partial[4][7 + i - j] += x;
      partial[5][3 - i / 2 + j] += x;
      partial[6][j] += x;
      partial[7][i / 2 + j] += x;
    }
  int32_t cost[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 8; i++) {
    cost[2] += partial[2][i] * partial[2][i];
    cost[6] += partial[6][i] * partial[6][i];
  }
  cost[2] *= DIV[8];
  cost[6] *= DIV[8];
#pragma unroll
  for (int i = 0; i < 7; i++) {
    cost[0] += (partial[0][i] * partial[0][i] + partial[0][14 - i] * partial[0][14 - i]) * DIV[i + 1];
    cost[4] += (partial[4][i] * partial[4][i] + partial[4][14 - i] * partial[4][14 - i]) * DIV[i + 1];
  }
  cost[0] += partial[0][7] * partial[0][7] * DIV[8];
  cost[4] += partial[4][7] * partial[4][7] * DIV[8];
#pragma unroll
  for (int i = 1; i < 8; i += 2) {
#pragma unroll
    for (int j = 0; j < 5; j++) cost[i] += partial[i][3 + j] * partial[i][3 + j];
    cost[i] *= DIV[8];
#pragma unroll
    for (int j = 0; j < 3; j++)
      cost[i] += (partial[i][j] * partial[i][j] + partial[i][10 - j] * partial[i][10 - j]) * DIV[2 * j + 2];
  }
  int32_t best = 0;
  int bd = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (cost[i] > best) { best = cost[i]; bd = i; }
  *var_out = (best - cost[(bd + 4) & 7]) >> 10;
  return bd;
}

// Neighbour differences of one pixel for one tap geometry (dir) and the
// clamp window of the filter output.
struct PixTaps {
  int32_t d_p[4];  // primary taps minus x: (k0 +, k0 -, k1 +, k1 -)
  int32_t d_s[8];  // secondary taps minus x: k0 (4), k1 (4)
  int32_t mn, mx;
};

__device__ __forceinline__ void gather_taps(const uint16_t* c, int dir, PixTaps& pt,
                                            const int (*dirs)[2] = kCdefDir) {
  const int32_t x = (int16_t)c[0];
  int32_t mn = x, mx = x;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int dk = dirs[dir + 2][k];
    const int32_t p0 = (int16_t)c[dk], p1 = (int16_t)c[-dk];
    pt.d_p[2 * k] = p0 - x;
    pt.d_p[2 * k + 1] = p1 - x;
    if ((uint16_t)p0 != kLarge) mx = max(mx, p0);
    if ((uint16_t)p1 != kLarge) mx = max(mx, p1);
    mn = min(mn, min(p0, p1));
    const int da = dirs[dir + 4][k], db = dirs[dir][k];
    const int32_t s[4] = {(int16_t)c[da], (int16_t)c[-da], (int16_t)c[db], (int16_t)c[-db]};
#pragma unroll
    for (int m = 0; m < 4; m++) {
      if ((uint16_t)s[m] != kLarge) mx = max(mx, s[m]);
      mn = min(mn, s[m]);
      pt.d_s[4 * k + m] = s[m] - x;
    }
  }
  pt.mn = mn;
  pt.mx = mx;
}

// Primary-tap sum for primary strength tq (> 0) with tap parity par.
__device__ __forceinline__ int32_t pri_sum(const PixTaps& t, int32_t tq, int par, int damping) {
  const int32_t sh = max(0, damping - msb32((uint32_t)tq));
  const int t0w = par ? 3 : 4, t1w = par ? 3 : 2;
  return t0w * (constrain(t.d_p[0], tq, sh) + constrain(t.d_p[1], tq, sh)) +
         t1w * (constrain(t.d_p[2], tq, sh) + constrain(t.d_p[3], tq, sh));
}

__device__ __forceinline__ int32_t cdef_out(int32_t x, int32_t sum, int32_t mn, int32_t mx) {
  const int32_t s16 = (int16_t)sum;  // the reference accumulates in wrapping i16
  return clampi(x + ((8 + s16 - (s16 < 0)) >> 4), mn, mx);
}

}  // namespace

namespace {
// is_8x8_all_skip (deblock.rs:652): every mi (up to 2 x 2) of the 8x8 at
// luma mi (mr, mc) carries the signaled skip flag.  skip == nullptr: none.
__device__ __forceinline__ bool all_skip_8x8(const uint8_t* skip, int mi_cols, int mi_rows, int mr, int mc) {
  if (!skip) return false;
  for (int r = mr; r < min(mr + 2, mi_rows); r++)
    for (int c = mc; c < min(mc + 2, mi_cols); c++)
      if (!skip[(long)r * mi_cols + c]) return false;
  return true;
}
}  // namespace

namespace {
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
  int32_t r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// constrain(d, thr, damping) with |d| = ad and sh = max(0, damping - msb(thr))
// wave-uniform: sign(d) * min(|d|, max(0, thr - (|d| >> sh))) as a clamp of d
// to +-max(0, thr - (|d| >> sh)).
__device__ __forceinline__ int32_t cons_u(int32_t d, uint32_t ad, uint32_t thr, uint32_t sh) {
  const int32_t m = (int32_t)__builtin_elementwise_sub_sat(thr, ad >> sh);
  return med3_i32(d, -m, m);
}
// x + round(sum / 16) (ties away from zero) clamped to [mn, mx], as the error
// s_v - out: sx = s_v - x, [lo, hi] = [mn - x, mx - x].  The tap sums stay
// far inside i16 (|sum| <= 2 * 4 * 60 + 2 * 2 * 60 + 12 * 16 = 912), so the
// reference's wrapping i16 accumulation never wraps.
__device__ __forceinline__ int32_t cdef_err(int32_t sum, int32_t sx, int32_t lo, int32_t hi) {
  return sx - med3_i32((sum + 8 + (sum >> 31)) >> 4, lo, hi);
}

// Differences of one pixel's taps from x for one direction (uniform), the
// clamp window relative to x, and the tap magnitudes.
struct TapsU {
  int32_t dp[4], ds[8];
  uint32_t ap[4], as[8];
  int32_t lo, hi;
};
// dirs: the 12 x 2 tap-offset table (kCdefDir, or its LDS copy when the
// direction differs between lanes)
__device__ __forceinline__ void gather_u(const uint16_t* c, int dir, int32_t x, TapsU& t,
                                         const int (*dirs)[2] = kCdefDir) {
  int32_t mn = x, mx = x;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int dk = dirs[dir + 2][k], da = dirs[dir + 4][k], db = dirs[dir][k];
    const int32_t v[6] = {(int16_t)c[dk], (int16_t)c[-dk], (int16_t)c[da], (int16_t)c[-da], (int16_t)c[db],
                          (int16_t)c[-db]};
#pragma unroll
    for (int m = 0; m < 6; m++) {
      mx = max(mx, (uint16_t)v[m] != kLarge ? v[m] : x);
      mn = min(mn, v[m]);
      const int32_t d = v[m] - x;
      if (m < 2) t.dp[2 * k + m] = d, t.ap[2 * k + m] = (uint32_t)abs(d);
      else t.ds[4 * k + m - 2] = d, t.as[4 * k + m - 2] = (uint32_t)abs(d);
    }
  }
  t.lo = mn - x;
  t.hi = mx - x;
}
// Secondary sums for S = 1..3 (strengths 1, 2, 4 << cs); out[0] = 0.
__device__ __forceinline__ void sec_u(const TapsU& t, int cs, int dmp, int32_t out[4]) {
  out[0] = 0;
#pragma unroll
  for (int S = 1; S < 4; S++) {
    const uint32_t st = (uint32_t)((S + (S == 3)) << cs);
    const uint32_t sh = (uint32_t)max(0, dmp - msb32(st));
    int32_t a = 0, b = 0;
#pragma unroll
    for (int m = 0; m < 4; m++) a += cons_u(t.ds[m], t.as[m], st, sh), b += cons_u(t.ds[4 + m], t.as[4 + m], st, sh);
    out[S] = 2 * a + b;
  }
}
}  // namespace

// One workgroup per 64x64 luma filter block: the 8x8 directions, then all
// 64 (primary, secondary) strength pairs for every pixel of the blocks in
// the fb's search list (not all-skip); per-pair SSE against the source
// reduced over the block.  dirvar (optional) receives (dir, var) per 8x8 for
// the chroma pass (0, 0 for blocks outside the list).
//
// A wave filters one 8x8 block at a time (lane = pixel), so the block's
// direction, variance and every adjusted primary strength are wave-uniform
// (scalar registers and scalar tap-offset loads).  Consecutive strengths P
// whose adjusted primary strength repeats give the same outputs, so their
// squared errors are computed once.  The per-pair sums of a wave's 16 blocks
// are reduced over each 16-lane row by DPP and added in LDS.
template <typename Pix>
__global__ __launch_bounds__(256) void k_cdef_luma(const Pix* __restrict__ rec, const Pix* __restrict__ src,
                                                   int width, int height, int bd, int damping, int sub, int nhfb,
                                                   const uint8_t* __restrict__ skip, int mi_cols, int mi_rows,
                                                   uint64_t* __restrict__ mse, int32_t* __restrict__ dirvar) {
  __shared__ uint16_t tile[70 * kTS];
  __shared__ int32_t sdir[64], svar[64];
  __shared__ uint8_t sact[64];
  __shared__ uint32_t tot[4][64];
  const int fb = xcd_chunked(blockIdx.x, gridDim.x);
  const int fbr = fb / nhfb, fbc = fb - fbr * nhfb;
  const int cs = bd - 8;
  const int y0 = fbr * 64, x0 = fbc * 64;
  const int vsize = min(64, height - y0), hsize = min(64, width - x0);
  const int tid = threadIdx.x;
  for (int i = tid; i < 70 * kTS; i += 256) {
    const int r = i / kTS, c = i - r * kTS;
    const int gy = y0 + r - kVB, gx = x0 + c - kHB;
    const bool in = (r < vsize + 2 * kVB) && (c < hsize + 2 * kHB) && gy >= 0 && gy < height && gx >= 0 &&
                    gx < width;
    tile[i] = in ? (uint16_t)rec[(long)gy * width + gx] : kLarge;
  }
  __syncthreads();
  const int nbx = hsize >> 3, nby = vsize >> 3;
  if (tid < 64) {
    const int by = tid >> 3, bx = tid & 7;
    int d = 0, v = 0;
    const bool act = by < nby && bx < nbx && !all_skip_8x8(skip, mi_cols, mi_rows, fbr * 16 + by * 2, fbc * 16 + bx * 2);
    if (act) d = find_dir(tile + (kVB + by * 8) * kTS + kHB + bx * 8, cs, &v);
    sdir[tid] = d;
    svar[tid] = v;
    sact[tid] = act;
    if (dirvar) {
      dirvar[((long)fb * 64 + tid) * 2] = d;
      dirvar[((long)fb * 64 + tid) * 2 + 1] = v;
    }
  }
  __syncthreads();

  const int dmp = damping + cs;
  const int wave = tid >> 6, lane = tid & 63, pr = lane >> 3, pc = lane & 7;
  uint32_t acc[64];
#pragma unroll
  for (int g = 0; g < 64; g++) acc[g] = 0;
  for (int blk = wave; blk < 64; blk += 4) {
    if (!__builtin_amdgcn_readfirstlane(sact[blk])) continue;
    const int dir = __builtin_amdgcn_readfirstlane(sdir[blk]);
    const int var = __builtin_amdgcn_readfirstlane(svar[blk]);
    const int r = (blk >> 3) * 8 + pr, c = (blk & 7) * 8 + pc;
    const uint16_t* cp = tile + (kVB + r) * kTS + kHB + c;
    const int32_t x = (int16_t)cp[0];
    TapsU tb, t0;
    gather_u(cp, dir, x, tb);
    gather_u(cp, 0, x, t0);
    int32_t sx = (int32_t)src[(long)(y0 + r) * width + x0 + c] - x;
    if ((pr % sub) != 0) sx = 0, tb.lo = tb.hi = t0.lo = t0.hi = 0;  // row not sampled: every error 0
    int32_t secb[4], sec0[4];
    sec_u(tb, cs, dmp, secb);
    sec_u(t0, cs, dmp, sec0);
    // P = 0 filters along direction 0 with no primary taps; (0, 0) is the
    // unfiltered block
    acc[0] = (uint32_t)mad_i24(sx, sx, (int32_t)acc[0]);
#pragma unroll
    for (int S = 1; S < 4; S++) {
      const int32_t e = cdef_err(sec0[S], sx, t0.lo, t0.hi);
      acc[S] = (uint32_t)mad_i24(e, e, (int32_t)acc[S]);
    }
    int tq_prev = -1;
    int32_t e2[4] = {0, 0, 0, 0};
#pragma unroll
    for (int P = 1; P < 16; P++) {
      const int tq = adjust_strength(P << cs, var);
      if (tq != tq_prev) {
        tq_prev = tq;
        int32_t pri = 0;
        if (tq) {
          const uint32_t sh = (uint32_t)max(0, dmp - msb32((uint32_t)tq));
          const int par = (tq >> cs) & 1, t0w = par ? 3 : 4, t1w = par ? 3 : 2;
          pri = mad_i24(t0w, cons_u(tb.dp[0], tb.ap[0], tq, sh) + cons_u(tb.dp[1], tb.ap[1], tq, sh),
                        __mul24(t1w, cons_u(tb.dp[2], tb.ap[2], tq, sh) + cons_u(tb.dp[3], tb.ap[3], tq, sh)));
        }
#pragma unroll
        for (int S = 0; S < 4; S++) {
          const int32_t e = cdef_err(pri + secb[S], sx, tb.lo, tb.hi);
          e2[S] = __mul24(e, e);
        }
      }
#pragma unroll
      for (int S = 0; S < 4; S++) acc[P * 4 + S] += (uint32_t)e2[S];
    }
  }
  // per-pair sums: 16-lane rows by DPP, the wave's four rows in scalars, one
  // LDS row per wave
#pragma unroll
  for (int g = 0; g < 64; g++) {
    const int32_t v = dpp_sum_i32<16>((int32_t)acc[g]);
    const uint32_t t = (uint32_t)__builtin_amdgcn_readlane(v, 0) + (uint32_t)__builtin_amdgcn_readlane(v, 16) +
                       (uint32_t)__builtin_amdgcn_readlane(v, 32) + (uint32_t)__builtin_amdgcn_readlane(v, 48);
    if (lane == 0) tot[wave][g] = t;
  }
  __syncthreads();
  if (tid < 64) {
    const uint64_t t = (uint64_t)tot[0][tid] + tot[1][tid] + tot[2][tid] + tot[3][tid];
    mse[(long)fb * 64 + tid] = (t >> (2 * cs)) * (uint64_t)sub;
  }
}

// Chroma (4:2:0): one workgroup per filter block for both chroma planes;
// 4x4 blocks with the co-located luma direction, unadjusted primary
// strength, damping - 1.  mse receives the U + V row (each plane's sum
// shifted on its own, as the reference does).
template <typename Pix>
__global__ __launch_bounds__(256) void k_cdef_chroma(const Pix* __restrict__ rec_u, const Pix* __restrict__ rec_v,
                                                     const Pix* __restrict__ src_u, const Pix* __restrict__ src_v,
                                                     int width, int height, int bd, int damping, int sub, int nhfb,
                                                     const uint8_t* __restrict__ skip, int mi_cols, int mi_rows,
                                                     const int32_t* __restrict__ dirvar,
                                                     uint64_t* __restrict__ mse) {
  __shared__ uint16_t tile[38 * kTS];
  __shared__ uint32_t red[4][64];
  __shared__ uint64_t total[64];
  __shared__ uint8_t sact[64];
  __shared__ int sdirs[12][2];  // per-lane directions: tap offsets from LDS, not per-lane constant loads
  const int fb = xcd_chunked(blockIdx.x, gridDim.x);
  const int fbr = fb / nhfb, fbc = fb - fbr * nhfb;
  const int cs = bd - 8, cw = width >> 1, ch = height >> 1;
  const int y0 = fbr * 32, x0 = fbc * 32;
  const int vsize = min(32, ch - y0), hsize = min(32, cw - x0);
  const int nr = vsize & ~3, nc = hsize & ~3;
  const int tid = threadIdx.x;
  const int dmp = damping + cs - 1;
  if (tid < 24) sdirs[tid >> 1][tid & 1] = kCdefDir[tid >> 1][tid & 1];
  if (tid < 64) {
    total[tid] = 0;
    sact[tid] = !all_skip_8x8(skip, mi_cols, mi_rows, fbr * 16 + (tid >> 3) * 2, fbc * 16 + (tid & 7) * 2);
  }
  for (int pl = 0; pl < 2; pl++) {
    const Pix* rec = pl ? rec_v : rec_u;
    const Pix* srcp = pl ? src_v : src_u;
    __syncthreads();  // previous plane's tile and red consumed
    for (int i = tid; i < 38 * kTS; i += 256) {
      const int r = i / kTS, c = i - r * kTS;
      const int gy = y0 + r - kVB, gx = x0 + c - kHB;
      const bool in = (r < vsize + 2 * kVB) && (c < hsize + 2 * kHB) && gy >= 0 && gy < ch && gx >= 0 && gx < cw;
      tile[i] = in ? (uint16_t)rec[(long)gy * cw + gx] : kLarge;
    }
    __syncthreads();
    uint32_t acc[64];
#pragma unroll
    for (int g = 0; g < 64; g++) acc[g] = 0;
    for (int p = tid; p < 1024; p += 256) {
      const int r = p >> 5, c = p & 31;
      if (r >= nr || c >= nc || !sact[(r >> 2) * 8 + (c >> 2)] || ((r & 3) % sub) != 0) continue;
      const int dir = dirvar[((long)fb * 64 + (r >> 2) * 8 + (c >> 2)) * 2];
      const uint16_t* cp = tile + (kVB + r) * kTS + kHB + c;
      const int32_t x = (int16_t)cp[0];
      TapsU tb, t0;
      gather_u(cp, dir, x, tb, sdirs);
      gather_u(cp, 0, x, t0);
      const int32_t sx = (int32_t)srcp[(long)(y0 + r) * cw + x0 + c] - x;
      int32_t secb[4], sec0[4];
      sec_u(tb, cs, dmp, secb);
      sec_u(t0, cs, dmp, sec0);
      acc[0] = (uint32_t)mad_i24(sx, sx, (int32_t)acc[0]);
#pragma unroll
      for (int S = 1; S < 4; S++) {
        const int32_t e = cdef_err(sec0[S], sx, t0.lo, t0.hi);
        acc[S] = (uint32_t)mad_i24(e, e, (int32_t)acc[S]);
      }
#pragma unroll
      for (int P = 1; P < 16; P++) {  // unadjusted primary strength P << cs, tap parity P & 1
        const uint32_t tq = (uint32_t)(P << cs), sh = (uint32_t)max(0, dmp - msb32(tq));
        const int t0w = (P & 1) ? 3 : 4, t1w = (P & 1) ? 3 : 2;
        const int32_t pri = t0w * (cons_u(tb.dp[0], tb.ap[0], tq, sh) + cons_u(tb.dp[1], tb.ap[1], tq, sh)) +
                            t1w * (cons_u(tb.dp[2], tb.ap[2], tq, sh) + cons_u(tb.dp[3], tb.ap[3], tq, sh));
#pragma unroll
        for (int S = 0; S < 4; S++) {
          const int32_t e = cdef_err(pri + secb[S], sx, tb.lo, tb.hi);
          acc[P * 4 + S] = (uint32_t)mad_i24(e, e, (int32_t)acc[P * 4 + S]);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < 64; g++) {
      const int32_t v = dpp_sum_i32<16>((int32_t)acc[g]);
      const uint32_t t = (uint32_t)__builtin_amdgcn_readlane(v, 0) + (uint32_t)__builtin_amdgcn_readlane(v, 16) +
                         (uint32_t)__builtin_amdgcn_readlane(v, 32) + (uint32_t)__builtin_amdgcn_readlane(v, 48);
      if ((tid & 63) == 0) red[tid >> 6][g] = t;
    }
    __syncthreads();
    if (tid < 64) {
      const uint64_t tot = (uint64_t)red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
      total[tid] += (tot >> (2 * cs)) * (uint64_t)sub;
    }
  }
  __syncthreads();
  if (tid < 64) mse[(long)fb * 64 + tid] = total[tid];
}

// The CDEF frame pass (apply_cdef_frame / _hbd, enc/cdef.rs:343 / :556),
// out of place: one workgroup per 64x64 filter block reads the pre-CDEF
// planes only, so filter blocks are independent.  The fb's luma tile (with
// the 3-row / 8-column border, kLarge outside the plane) is staged in LDS,
// the 8x8 directions are found, luma filters, then each chroma plane's tile
// replaces it and its 4x4 blocks filter along the luma directions.  Every
// pixel of the fb is written (the input where nothing filters).
namespace {
__device__ __forceinline__ int32_t cdef_pixel(const uint16_t* cp, int pri, int sec, int dir, int dmp, int cs,
                                              const int (*dirs)[2]) {
  PixTaps pt;
  gather_taps(cp, dir, pt, dirs);
  int32_t sum = pri ? pri_sum(pt, pri, (pri >> cs) & 1, dmp) : 0;
  if (sec) {
    const int32_t sh = max(0, dmp - msb32((uint32_t)sec));
#pragma unroll
    for (int m = 0; m < 8; m++) sum += (m < 4 ? 2 : 1) * constrain(pt.d_s[m], sec, sh);
  }
  return cdef_out((int16_t)cp[0], sum, pt.mn, pt.mx);
}
}  // namespace

template <typename Pix>
__global__ __launch_bounds__(256) void k_cdef_apply(const Pix* __restrict__ in_y, const Pix* __restrict__ in_u,
                                                    const Pix* __restrict__ in_v, Pix* __restrict__ out_y,
                                                    Pix* __restrict__ out_u, Pix* __restrict__ out_v, int width,
                                                    int height, int bd, int damping, int nhfb,
                                                    const uint8_t* __restrict__ fb_str,
                                                    const uint8_t* __restrict__ skip, int mi_cols, int mi_rows,
                                                    unsigned long long* __restrict__ stats) {
  __shared__ uint16_t tile[70 * kTS];
  __shared__ int32_t sdir[64], svar[64];
  __shared__ uint8_t sact[64];
  __shared__ int sdirs[12][2];
  __shared__ unsigned long long red[2];
  const int fb = xcd_chunked(blockIdx.x, gridDim.x);
  const int fbr = fb / nhfb, fbc = fb - fbr * nhfb;
  const int cs = bd - 8;
  const int y0 = fbr * 64, x0 = fbc * 64;
  const int vsize = min(64, height - y0), hsize = min(64, width - x0);
  const int tid = threadIdx.x;
  for (int i = tid; i < 70 * kTS; i += 256) {
    const int r = i / kTS, c = i - r * kTS;
    const int gy = y0 + r - kVB, gx = x0 + c - kHB;
    const bool in = (r < vsize + 2 * kVB) && (c < hsize + 2 * kHB) && gy >= 0 && gy < height && gx >= 0 &&
                    gx < width;
    tile[i] = in ? (uint16_t)in_y[(long)gy * width + gx] : kLarge;
  }
  if (tid < 24) sdirs[tid >> 1][tid & 1] = kCdefDir[tid >> 1][tid & 1];
  if (tid < 2) red[tid] = 0;
  __syncthreads();
  const int nbx = hsize >> 3, nby = vsize >> 3;
  if (tid < 64) {
    const int by = tid >> 3, bx = tid & 7;
    int d = 0, v = 0;
    const bool act = by < nby && bx < nbx && !all_skip_8x8(skip, mi_cols, mi_rows, fbr * 16 + by * 2, fbc * 16 + bx * 2);
    if (act) d = find_dir(tile + (kVB + by * 8) * kTS + kHB + bx * 8, cs, &v);
    sdir[tid] = d;
    svar[tid] = v;
    sact[tid] = act;
  }
  __syncthreads();
  // the fb's strengths: primary idx / 4, secondary idx % 4 (3 -> 4), << cs
  const int sy = fb_str[2 * fb], suv = fb_str[2 * fb + 1];
  const int lvl_y = (sy >> 2) << cs, sec_y = ((sy & 3) + ((sy & 3) == 3)) << cs;
  const int lvl_uv = in_u ? (suv >> 2) << cs : 0, sec_uv = in_u ? ((suv & 3) + ((suv & 3) == 3)) << cs : 0;
  const int dmp = damping + cs;
  uint32_t n_f = 0, n_c = 0;
  for (int p = tid; p < 4096; p += 256) {
    const int r = p >> 6, c = p & 63;
    if (r >= vsize || c >= hsize) continue;
    const int blk = (r >> 3) * 8 + (c >> 3);
    const uint16_t* cp = tile + (kVB + r) * kTS + kHB + c;
    int32_t v = cp[0];
    if ((lvl_y || sec_y) && sact[blk]) {
      const int t = adjust_strength(lvl_y, svar[blk]);
      if (t || sec_y) {
        const int32_t f = cdef_pixel(cp, t, sec_y, lvl_y ? sdir[blk] : 0, dmp, cs, sdirs);
        n_f++;
        n_c += f != v;
        v = f;
      }
    }
    out_y[(long)(y0 + r) * width + x0 + c] = (Pix)v;
  }
  if (in_u) {
    const int cw = width >> 1, chh = height >> 1, cy0 = fbr * 32, cx0 = fbc * 32;
    const int cv = vsize >> 1, chs = hsize >> 1;
    for (int pl = 0; pl < 2; pl++) {
      const Pix* in = pl ? in_v : in_u;
      Pix* out = pl ? out_v : out_u;
      __syncthreads();  // the previous tile is consumed
      for (int i = tid; i < 38 * kTS; i += 256) {
        const int r = i / kTS, c = i - r * kTS;
        const int gy = cy0 + r - kVB, gx = cx0 + c - kHB;
        const bool ok = (r < cv + 2 * kVB) && (c < chs + 2 * kHB) && gy >= 0 && gy < chh && gx >= 0 && gx < cw;
        tile[i] = ok ? (uint16_t)in[(long)gy * cw + gx] : kLarge;
      }
      __syncthreads();
      for (int p = tid; p < 1024; p += 256) {
        const int r = p >> 5, c = p & 31;
        if (r >= cv || c >= chs) continue;
        const int blk = (r >> 2) * 8 + (c >> 2);
        const uint16_t* cp = tile + (kVB + r) * kTS + kHB + c;
        int32_t v = cp[0];
        if ((lvl_uv || sec_uv) && sact[blk]) {
          const int32_t f = cdef_pixel(cp, lvl_uv, sec_uv, lvl_uv ? sdir[blk] : 0, dmp - 1, cs, sdirs);
          n_f++;
          n_c += f != v;
          v = f;
        }
        out[(long)(cy0 + r) * cw + cx0 + c] = (Pix)v;
      }
    }
  }
  if (stats) {
    uint64_t a = n_f, b = n_c;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m, 64), b += __shfl_xor(b, m, 64);
    if ((tid & 63) == 0) {
      atomicAdd(&red[0], (unsigned long long)a);
      atomicAdd(&red[1], (unsigned long long)b);
    }
    __syncthreads();
    if (tid == 0) {
      atomicAdd(&stats[0], red[0]);
      atomicAdd(&stats[1], red[1]);
    }
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t cdef_apply(const void* in_y, const void* in_u, const void* in_v, void* out_y, void* out_u, void* out_v,
                      int width, int height, int bd, int damping, const uint8_t* fb_str, const uint8_t* skip,
                      int mi_cols, int mi_rows, uint64_t* stats, hipStream_t st) {
  const int nhfb = (width + 63) / 64, nvfb = (height + 63) / 64;
  auto* s = reinterpret_cast<unsigned long long*>(stats);
  if (s) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, s, 2);
  const dim3 grid((unsigned)(nhfb * nvfb));
  if (bd == 8)
    hipLaunchKernelGGL((k_cdef_apply<uint8_t>), grid, dim3(256), 0, st, (const uint8_t*)in_y, (const uint8_t*)in_u,
                       (const uint8_t*)in_v, (uint8_t*)out_y, (uint8_t*)out_u, (uint8_t*)out_v, width, height, bd,
                       damping, nhfb, fb_str, skip, mi_cols, mi_rows, s);
  else
    hipLaunchKernelGGL((k_cdef_apply<uint16_t>), grid, dim3(256), 0, st, (const uint16_t*)in_y,
                       (const uint16_t*)in_u, (const uint16_t*)in_v, (uint16_t*)out_y, (uint16_t*)out_u,
                       (uint16_t*)out_v, width, height, bd, damping, nhfb, fb_str, skip, mi_cols, mi_rows, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Restoration units
// ---------------------------------------------------------------------------

namespace {
// foreach_rest_unit_in_tile (restoration.rs:1395) for a whole plane.
__device__ __forceinline__ void unit_limits(int unit_size, int pw, int ph, int ss_y, int ux, int uy, int& hs,
                                            int& he, int& vs, int& ve) {
  const int ext = unit_size * 3 / 2;
  const int y0 = uy * unit_size, x0 = ux * unit_size;
  const int h = (ph - y0 < ext) ? ph - y0 : unit_size;
  const int w = (pw - x0 < ext) ? pw - x0 : unit_size;
  const int voff = 8 >> ss_y;
  vs = max(0, y0 - voff);
  ve = y0 + h;
  if (ve < ph) ve -= voff;
  hs = x0;
  he = x0 + w;
}

// Workgroup-wide u64 sum (256 threads); every thread gets the result.
__device__ __forceinline__ uint64_t wg_sum_u64(uint64_t v, uint64_t* red) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}
}  // namespace

// Wiener statistics (compute_stats / compute_stats_hbd) on the matrix cores.
//
// With y_p the WIN x WIN window of mean-removed dgd samples around pixel p
// (entry k = dx * WIN + dy) and x_p its mean-removed source sample,
// H = sum_p y_p y_p^T and M = sum_p y_p x_p: one Gram matrix Z^T Z over the
// unit's pixels, Z's row p holding y_p and x_p.  Each |value| <= 1023 is
// split into two int8 digits, v = 32 * vh + vl (vh in [-32, 31], vl in
// [0, 31]), so Z has 2 * WIN^2 + 2 columns (yh, yl, xh, xl) and Z^T Z is
// accumulated exactly in i32 by v_mfma_i32_16x16x64_i8 over 16-column
// blocks, upper-triangle block pairs only (28 for WIN = 7, 10 for WIN = 5);
// H = 1024 HH + 32 (HL + HL^T) + LL and M likewise, in i64, at the end.
//
// A workgroup (four waves) owns one unit (or one explicit window) and walks
// it in tiles of 16 rows x 128 columns.  A tile is staged in LDS as digit
// planes, one per (digit, dx) with its columns shifted by dx, so that the
// 16 samples of feature (dx, dy) at 16 consecutive pixels of a row are one
// aligned 16-byte read; pixel columns past the tile read as 0 in every
// plane.  A K step is 4 rows x 16 columns (64 pixels): lane group g (lanes
// 16g .. 16g + 15) covers row g, and a lane's fragment of block b (feature
// 16b + lane % 16) is at once its A operand (row of Z^T) and its B operand
// (column of Z), both MFMA operands taking K from the lane group and byte in
// the same way.  Wave w takes row quad w of the tile.  The four waves'
// accumulators meet in LDS (ds_add) and the final H / M are formed from it.
// The i32 sums stay exact: |digit products| <= 1024 over < 2^15 pixels.
constexpr int kWnRB = 16, kWnCT = 128;
// LDS row stride of the digit planes: 9 x 16 bytes, so the seven dy rows a
// lane group reads start 36 dwords apart and spread over the banks
constexpr int kWnRS = kWnCT + 16;
typedef int32_t wn_v4i __attribute__((ext_vector_type(4)));

template <int WIN>
struct WnGeom {
  static constexpr int W2 = WIN * WIN, HALF = WIN / 2, NF = 2 * W2 + 2, NB = (NF + 15) / 16;
  static constexpr int NT = NB * (NB + 1) / 2;
  // planes 16 bytes longer than their rows: the dx planes of a lane group
  // then land on different banks as well (two-way conflicts at most)
  static constexpr int PR = kWnRB + 2 * HALF, PLANE = PR * kWnRS + 16, XOFF = 2 * WIN * PLANE;
  static constexpr int LDS = XOFF + 2 * kWnRB * kWnRS;
  static_assert(NT * 256 * 4 <= LDS, "reduction fits the staging area");
};

namespace {
// bytes [0, vb) of a 16-byte chunk kept, the rest cleared
__device__ __forceinline__ wn_v4i keep_bytes(wn_v4i v, int vb) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int n = vb - 4 * q;
    const uint32_t m = n >= 4 ? 0xffffffffu : (n <= 0 ? 0u : (1u << (8 * n)) - 1u);
    v[q] &= (int32_t)m;
  }
  return v;
}
}  // namespace

template <typename Pix, int WIN>
__global__ __launch_bounds__(256) void k_wiener_stats(const Pix* __restrict__ dgd, const Pix* __restrict__ src,
                                                      int pw, int ph, int unit_size, int ss_y, int nh, int4 lim,
                                                      int64_t* __restrict__ out_m, int64_t* __restrict__ out_h) {
  using G = WnGeom<WIN>;
  constexpr int W2 = G::W2, HALF = G::HALF, NB = G::NB;
  constexpr int64_t DIV = sizeof(Pix) == 1 ? 1 : 4;  // compute_stats_hbd scales by 1/4
  __shared__ __align__(16) uint8_t lds[G::LDS];
  __shared__ uint64_t red[4];
  const int u = xcd_chunked(blockIdx.x, gridDim.x), uy = u / nh, ux = u - uy * nh;
  int hs, he, vs, ve;
  if (lim.y > lim.x) {  // one window given explicitly (the plain compute_stats)
    hs = lim.x; he = lim.y; vs = lim.z; ve = lim.w;
  } else {
    unit_limits(unit_size, pw, ph, ss_y, ux, uy, hs, he, vs, ve);
  }
  const int uw = he - hs, uh = ve - vs;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4;
  uint64_t s = 0;
  for (int r = w; r < uh; r += 4)
    for (int c = lane; c < uw; c += 64) s += dgd[(long)(vs + r) * pw + hs + c];
  const uint64_t area = (uint64_t)uw * (uint64_t)uh;
  const int32_t avg = (int32_t)(area ? wg_sum_u64(s, red) / area : 0);

  // LDS offset of this lane's feature in block b (row g of a K step); the
  // padding features past the last real one read feature 0 (never used)
  int addr[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    int f = 16 * b + (lane & 15);
    if (f >= G::NF) f = 0;
    if (f < 2 * W2) {
      const int d = f >= W2, k = f - d * W2, dx = k / WIN, dy = k - dx * WIN;
      addr[b] = (d * WIN + dx) * G::PLANE + (dy + g) * kWnRS;
    } else {
      addr[b] = G::XOFF + (f - 2 * W2) * kWnRB * kWnRS + g * kWnRS;
    }
  }
  wn_v4i acc[G::NT];
#pragma unroll
  for (int t = 0; t < G::NT; t++) acc[t] = wn_v4i{0, 0, 0, 0};

  for (int r0 = 0; r0 < uh; r0 += kWnRB) {
    const int br = min(kWnRB, uh - r0);
    for (int ct0 = 0; ct0 < uw; ct0 += kWnCT) {
      const int tw = min(kWnCT, uw - ct0), nc = (tw + 15) >> 4;
      __syncthreads();  // previous tile's fragments read
      const int ny = G::PR * nc, nx = br * nc;
      for (int task = tid; task < ny + nx; task += 256) {
        if (task < ny) {
          // plane row rr, chunk j: samples at columns 16j - HALF .. 16j + 15 + HALF
          const int rr = task / nc, j = task - rr * nc;
          const Pix* row = dgd + (long)clampi(vs + r0 - HALF + rr, 0, ph - 1) * pw;
          const int cb = hs + ct0 + 16 * j - HALF;
          uint32_t wh[6], wl[6];
#pragma unroll
          for (int q = 0; q < 6; q++) {
            uint32_t h = 0, l = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
              const int idx = 4 * q + e;
              if (idx < 16 + 2 * HALF) {
                const int32_t y = (int32_t)row[clampi(cb + idx, 0, pw - 1)] - avg;
                h |= (uint32_t)(uint8_t)(y >> 5) << (8 * e);
                l |= (uint32_t)(y & 31) << (8 * e);
              }
            }
            wh[q] = h, wl[q] = l;
          }
          const int vb = tw - 16 * j;
#pragma unroll
          for (int dx = 0; dx < WIN; dx++) {
            wn_v4i vh, vl;
#pragma unroll
            for (int q = 0; q < 4; q++) {
              const int n = q + (dx >> 2), sh = dx & 3;
              vh[q] = (int32_t)(sh ? __builtin_amdgcn_alignbyte(wh[n + 1], wh[n], sh) : wh[n]);
              vl[q] = (int32_t)(sh ? __builtin_amdgcn_alignbyte(wl[n + 1], wl[n], sh) : wl[n]);
            }
            if (vb < 16) vh = keep_bytes(vh, vb), vl = keep_bytes(vl, vb);
            *(wn_v4i*)(lds + dx * G::PLANE + rr * kWnRS + 16 * j) = vh;
            *(wn_v4i*)(lds + (WIN + dx) * G::PLANE + rr * kWnRS + 16 * j) = vl;
          }
        } else {
          const int t2 = task - ny, rr = t2 / nc, j = t2 - rr * nc;
          const Pix* row = src + (long)(vs + r0 + rr) * pw;
          const int cb = hs + ct0 + 16 * j;
          wn_v4i vh, vl;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            uint32_t h = 0, l = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
              const int32_t x = (int32_t)row[min(cb + 4 * q + e, pw - 1)] - avg;
              h |= (uint32_t)(uint8_t)(x >> 5) << (8 * e);
              l |= (uint32_t)(x & 31) << (8 * e);
            }
            vh[q] = (int32_t)h, vl[q] = (int32_t)l;
          }
          const int vb = tw - 16 * j;
          if (vb < 16) vh = keep_bytes(vh, vb), vl = keep_bytes(vl, vb);
          *(wn_v4i*)(lds + G::XOFF + rr * kWnRS + 16 * j) = vh;
          *(wn_v4i*)(lds + G::XOFF + kWnRB * kWnRS + rr * kWnRS + 16 * j) = vl;
        }
      }
      __syncthreads();
      if (4 * w < br) {  // wave w: tile rows 4w .. 4w + 3
        const bool rowok = 4 * w + g < br, ragged = 4 * w + 4 > br;
        for (int j = 0; j < nc; j++) {
          const int off = 4 * w * kWnRS + 16 * j;
          wn_v4i fr[NB];
#pragma unroll
          for (int b = 0; b < NB; b++) fr[b] = *(const wn_v4i*)(lds + addr[b] + off);
          if (ragged) {
#pragma unroll
            for (int b = 0; b < NB; b++)
              if (!rowok) fr[b] = wn_v4i{0, 0, 0, 0};
          }
          int t = 0;
#pragma unroll
          for (int bi = 0; bi < NB; bi++)
#pragma unroll
            for (int bj = bi; bj < NB; bj++, t++)
              acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fr[bi], fr[bj], acc[t], 0, 0, 0);
        }
      }
    }
  }
  // the four waves' blocks summed in LDS: block t, row i = 4g + r, column lane % 16
  __syncthreads();
  int32_t* R = (int32_t*)lds;
  for (int i = tid; i < G::NT * 256; i += 256) R[i] = 0;
  __syncthreads();
#pragma unroll
  for (int t = 0; t < G::NT; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) atomicAdd(&R[t * 256 + (4 * g + r) * 16 + (lane & 15)], acc[t][r]);
  __syncthreads();
  auto at = [&](int i, int j) -> int64_t {  // (Z^T Z)[i][j] from the upper block triangle
    if ((i >> 4) > (j >> 4)) {
      const int x = i;
      i = j, j = x;
    }
    const int bi = i >> 4, bj = j >> 4, t = bi * NB - bi * (bi - 1) / 2 + (bj - bi);
    return R[t * 256 + (i & 15) * 16 + (j & 15)];
  };
  auto join = [&](int i, int j) -> int64_t {  // digits (h, l) of i and j recombined
    return 1024 * at(i, j) + 32 * (at(i, W2 + j) + at(j, W2 + i)) + at(W2 + i, W2 + j);
  };
  for (int i = tid; i < W2 * W2 + W2; i += 256) {
    if (i < W2 * W2) {
      const int k = i / W2, l = i - k * W2;
      out_h[(long)u * W2 * W2 + i] = join(k, l) / DIV;
    } else {
      const int k = i - W2 * W2, x = 2 * W2;
      const int64_t m = 1024 * at(k, x) + 32 * (at(k, x + 1) + at(W2 + k, x)) + at(W2 + k, x + 1);
      out_m[(long)u * W2 + k] = m / DIV;
    }
  }
}

// Per-unit SSE of two planes over the unit limits.
template <typename Pix>
__global__ __launch_bounds__(256) void k_unit_sse(const Pix* __restrict__ a, const Pix* __restrict__ b, int pw,
                                                  int ph, int unit_size, int ss_y, int nh,
                                                  uint64_t* __restrict__ out) {
  __shared__ uint64_t red[4];
  const int u = xcd_chunked(blockIdx.x, gridDim.x), uy = u / nh, ux = u - uy * nh;
  int hs, he, vs, ve;
  unit_limits(unit_size, pw, ph, ss_y, ux, uy, hs, he, vs, ve);
  const int uw = he - hs, uh = ve - vs, n = uw * uh, dw = uw % 16, dh = uh % 16;
  // sse_region_hbd's parts (each truncated to 32 bits): the right strip of
  // uw % 16 columns, the bottom strip left of it, the 16x16 blocks
  uint64_t sb = 0, sr = 0, st = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int r = i / uw, c = i - r * uw;
    const long o = (long)(vs + r) * pw + hs + c;
    const int64_t e = (int64_t)a[o] - b[o];
    const uint64_t e2 = (uint64_t)(e * e);
    if (sizeof(Pix) == 1 || (c < uw - dw && r < uh - dh)) sb += e2;
    else if (c >= uw - dw) sr += e2;
    else st += e2;
  }
  uint64_t t = wg_sum_u64(sb, red);
  if (sizeof(Pix) == 2) t += (uint32_t)wg_sum_u64(sr, red) + (uint64_t)(uint32_t)wg_sum_u64(st, red);
  if (threadIdx.x == 0) out[u] = t;
}

// ---------------------------------------------------------------------------
// Wiener unit filter (loop_restoration_filter_unit without stripe
// boundaries, restoration.rs:1288 / :1921) and the finer tap search
// (finer_tile_search_wiener, enc/restoration.rs:513)
// ---------------------------------------------------------------------------
// Per output pixel the unit filter is wiener_convolve_add_src (:138 / :1820)
// on the edge-extended plane: the stripe and processing-unit split of the
// reference changes no pixel of the unit.  A workgroup filters a tile of at
// most 64 x 64 pixels of one unit: the input with a 3-pixel border (edge
// replicated) into LDS, the horizontal pass over the tile's rows + 6 into
// LDS (round 3, clamp to [0, 2^(bd + 5) - 1], fits u16), then the vertical
// pass (round 11, clamp to the pixel range).  The 8th taps are 0 and are
// not read.
struct WienerUnitDev {
  int16_t hfilter[8], vfilter[8];
  uint8_t type, pad[7];
};
static_assert(sizeof(WienerUnitDev) == 40, "zav1_wiener_unit layout");
constexpr int kWtS = 72;  // input tile stride (64 + 6, rounded up)

template <typename Pix, class Emit>
__device__ __forceinline__ void wiener_tile(const Pix* __restrict__ dgd, int pw, int ph, int bd, int tx0, int ty0,
                                            int tw, int th, const int* hf, const int* vf, uint16_t* in,
                                            uint16_t* mid, Emit&& emit) {
  const int tid = threadIdx.x, ir = th + 6, ic = tw + 6;
  for (int i = tid; i < ir * ic; i += 256) {
    const int r = i / ic, c = i - r * ic;
    const int gy = min(max(ty0 + r - 3, 0), ph - 1), gx = min(max(tx0 + c - 3, 0), pw - 1);
    in[r * kWtS + c] = (uint16_t)dgd[(long)gy * pw + gx];
  }
  __syncthreads();
  const int clamp_h = (1 << (bd + 5)) - 1;
  for (int i = tid; i < ir * tw; i += 256) {
    const int r = i / tw, c = i - r * tw;
    const uint16_t* q = in + r * kWtS + c;
    int32_t sum = ((int32_t)q[3] << 7) + (1 << (bd + 6));
#pragma unroll
    for (int k = 0; k < 7; k++) sum += (int32_t)q[k] * hf[k];
    mid[r * 64 + c] = (uint16_t)min(max((sum + 4) >> 3, 0), clamp_h);
  }
  __syncthreads();
  const int mx = (1 << bd) - 1;
  for (int i = tid; i < th * tw; i += 256) {
    const int r = i / tw, c = i - r * tw;
    const uint16_t* q = mid + r * 64 + c;
    int32_t sum = ((int32_t)q[3 * 64] << 7) - (1 << (bd + 10));
#pragma unroll
    for (int k = 0; k < 7; k++) sum += (int32_t)q[k * 64] * vf[k];
    emit(r, c, min(max((sum + 1024) >> 11, 0), mx));
  }
  __syncthreads();
}

// The plane by tiles of T = min(64, unit size) aligned to the unit grid
// (rows from 0, then T k - (8 >> ss_y)), so each tile lies in one unit.
template <typename Pix>
__global__ __launch_bounds__(256) void k_wiener_apply(const Pix* __restrict__ dgd, Pix* __restrict__ out, int pw,
                                                      int ph, int bd, int unit_size, int ss_y, int nh, int nv, int T,
                                                      int ntc, const WienerUnitDev* __restrict__ units) {
  __shared__ uint16_t in[70 * kWtS], mid[70 * 64];
  const int t = blockIdx.x, tr = t / ntc, tc = t - tr * ntc, voff = 8 >> ss_y;
  const int ty0 = tr == 0 ? 0 : tr * T - voff, ty1 = min((tr + 1) * T - voff, ph);
  if (ty0 >= ph) return;
  const int tx0 = tc * T, tw = min(T, pw - tx0), th = ty1 - ty0;
  const int ux = min(tx0 / unit_size, nh - 1), uy = min((ty0 + voff) / unit_size, nv - 1);
  const WienerUnitDev& wu = units[uy * nh + ux];
  if (wu.type == 0) {
    for (int i = threadIdx.x; i < tw * th; i += 256) {
      const long o = (long)(ty0 + i / tw) * pw + tx0 + i % tw;
      out[o] = dgd[o];
    }
    return;
  }
  int hf[7], vf[7];
#pragma unroll
  for (int k = 0; k < 7; k++) hf[k] = wu.hfilter[k], vf[k] = wu.vfilter[k];
  wiener_tile(dgd, pw, ph, bd, tx0, ty0, tw, th, hf, vf, in, mid,
              [&](int r, int c, int v) { out[(long)(ty0 + r) * pw + tx0 + c] = (Pix)v; });
}

// One trial of the search: the unit [hs, he) x [vs, ve) filtered with taps
// (LDS, the workgroup's current filter) and its sse_region against src
// (the 10-bit parts truncated as the reference does).  Every thread gets
// the result.
template <typename Pix>
__device__ __noinline__ uint64_t wiener_trial(const Pix* __restrict__ dgd, const Pix* __restrict__ src, int pw,
                                              int ph, int bd, int hs, int he, int vs, int ve, const int* taps) {
  __shared__ uint16_t in[70 * kWtS], mid[70 * 64];
  __shared__ uint64_t red[4];
  int hf[7], vf[7];
#pragma unroll
  for (int k = 0; k < 7; k++) hf[k] = taps[k], vf[k] = taps[8 + k];
  const int uw = he - hs, uh = ve - vs, dw = uw % 16, dh = uh % 16;
  uint64_t sb = 0, sr = 0, st = 0;
  for (int ty = vs; ty < ve; ty += 64)
    for (int tx = hs; tx < he; tx += 64)
      wiener_tile(dgd, pw, ph, bd, tx, ty, min(64, he - tx), min(64, ve - ty), hf, vf, in, mid,
                  [&](int r, int c, int v) {
                    const int64_t e = (int64_t)src[(long)(ty + r) * pw + tx + c] - v;
                    const uint64_t e2 = (uint64_t)(e * e);
                    const int x = tx + c - hs, y = ty + r - vs;
                    if (sizeof(Pix) == 1 || (x < uw - dw && y < uh - dh)) sb += e2;
                    else if (x >= uw - dw) sr += e2;
                    else st += e2;
                  });
  uint64_t t = wg_sum_u64(sb, red);
  if (sizeof(Pix) == 2) t += (uint32_t)wg_sum_u64(sr, red) + (uint64_t)(uint32_t)wg_sum_u64(st, red);
  return t;
}

// One workgroup per unit runs the unit's whole hill climb; the decisions
// are workgroup-uniform (every thread sees each trial's SSE), thread 0
// moves the taps in LDS between trials.
template <typename Pix>
__global__ __launch_bounds__(256) void k_wiener_refine(const Pix* __restrict__ dgd, const Pix* __restrict__ src,
                                                       int pw, int ph, int bd, int win, int unit_size, int ss_y,
                                                       int nh, int use_refinement, int max_one_step,
                                                       WienerUnitDev* __restrict__ units,
                                                       int64_t* __restrict__ err_out) {
  __shared__ int taps[16];
  const int u = blockIdx.x, uy = u / nh, ux = u - uy * nh, tid = threadIdx.x;
  int hs, he, vs, ve;
  unit_limits(unit_size, pw, ph, ss_y, ux, uy, hs, he, vs, ve);
  if (tid < 8) taps[tid] = units[u].hfilter[tid], taps[8 + tid] = units[u].vfilter[tid];
  __syncthreads();
  int64_t err = (int64_t)wiener_trial(dgd, src, pw, ph, bd, hs, he, vs, ve, taps);
  constexpr int tmin[3] = {3 - 8, -7 - 16, 15 - 32}, tmax[3] = {3 - 1 + 8, -7 - 1 + 16, 15 - 1 + 32};
  auto move = [&](int f, int p, int d) {
    __syncthreads();
    if (tid == 0) taps[f * 8 + p] += d, taps[f * 8 + 6 - p] += d, taps[f * 8 + 3] -= 2 * d;
    __syncthreads();
  };
  if (use_refinement) {
    const int start = 4, end = max_one_step ? 4 : 1, plane_off = (7 - win) >> 1;
    for (int s = start; s >= end; s >>= 1)
      for (int f = 0; f < 2; f++)
        for (int p = plane_off; p < 3; p++) {
          bool skip = false;
          for (;;) {  // minus
            if (taps[f * 8 + p] - s >= tmin[p]) {
              move(f, p, -s);
              const int64_t e2 = (int64_t)wiener_trial(dgd, src, pw, ph, bd, hs, he, vs, ve, taps);
              if (e2 > err) {
                move(f, p, s);
              } else {
                err = e2;
                skip = true;
                if (s == start && !max_one_step) continue;
              }
            }
            break;
          }
          if (skip) break;  // a kept minus move ends this pass over the taps
          for (;;) {        // plus
            if (taps[f * 8 + p] + s <= tmax[p]) {
              move(f, p, s);
              const int64_t e2 = (int64_t)wiener_trial(dgd, src, pw, ph, bd, hs, he, vs, ve, taps);
              if (e2 > err) {
                move(f, p, -s);
              } else {
                err = e2;
                if (s == start && !max_one_step) continue;
              }
            }
            break;
          }
        }
  }
  __syncthreads();
  if (tid < 8) units[u].hfilter[tid] = (int16_t)taps[tid], units[u].vfilter[tid] = (int16_t)taps[8 + tid];
  if (tid == 0) err_out[u] = err;
}

hipError_t wiener_apply(const void* dgd, void* out, int pw, int ph, int bd, int unit_size, int ss_y, int nh, int nv,
                        const void* units, hipStream_t st) {
  const int T = unit_size < 64 ? unit_size : 64, voff = 8 >> ss_y;
  const int ntc = (pw + T - 1) / T, ntr = (ph + voff + T - 1) / T;
  const dim3 grid((unsigned)(ntc * ntr));
  const auto* wu = static_cast<const WienerUnitDev*>(units);
  if (bd == 8)
    hipLaunchKernelGGL((k_wiener_apply<uint8_t>), grid, dim3(256), 0, st, (const uint8_t*)dgd, (uint8_t*)out, pw, ph,
                       bd, unit_size, ss_y, nh, nv, T, ntc, wu);
  else
    hipLaunchKernelGGL((k_wiener_apply<uint16_t>), grid, dim3(256), 0, st, (const uint16_t*)dgd, (uint16_t*)out, pw,
                       ph, bd, unit_size, ss_y, nh, nv, T, ntc, wu);
  return hipGetLastError();
}

hipError_t wiener_refine(const void* dgd, const void* src, int pw, int ph, int bd, int win, int unit_size, int ss_y,
                         int nh, int nv, int use_refinement, int max_one_step, void* units, int64_t* err,
                         hipStream_t st) {
  auto* wu = static_cast<WienerUnitDev*>(units);
  const dim3 grid((unsigned)(nh * nv));
  if (bd == 8)
    hipLaunchKernelGGL((k_wiener_refine<uint8_t>), grid, dim3(256), 0, st, (const uint8_t*)dgd, (const uint8_t*)src,
                       pw, ph, bd, win, unit_size, ss_y, nh, use_refinement, max_one_step, wu, err);
  else
    hipLaunchKernelGGL((k_wiener_refine<uint16_t>), grid, dim3(256), 0, st, (const uint16_t*)dgd,
                       (const uint16_t*)src, pw, ph, bd, win, unit_size, ss_y, nh, use_refinement, max_one_step, wu,
                       err);
  return hipGetLastError();
}

// Self-guided parameter sets {r0, s0, r1, s1} (AV1 Sgr_Params).
__constant__ int kSgrParams[16][4] = {
    {2, 140, 1, 3236}, {2, 112, 1, 2158}, {2, 93, 1, 1618}, {2, 80, 1, 1438}, {2, 70, 1, 1295}, {2, 58, 1, 1177},
    {2, 47, 1, 1079},  {2, 37, 1, 996},   {2, 30, 1, 925},  {2, 25, 1, 863},  {0, -1, 1, 2589}, {0, -1, 1, 1618},
    {0, -1, 1, 1177},  {0, -1, 1, 925},   {2, 56, 0, -1},   {2, 22, 0, -1}};

namespace {
__device__ __forceinline__ int x_by_xplus1(int z) {
  return z == 0 ? 1 : (z >= 255 ? 256 : (256 * z + (z + 1) / 2) / (z + 1));
}
}  // namespace

// Self-guided projection statistics, strip form.
//
// One wave walks one half of a restoration-unit row (an even number of
// rows, then the rest) of one 64-column strip of the
// plane (lane L <-> column x0 - 3 + L, clamped to the plane like every
// other read), carrying four parameter sets; a workgroup is the four waves
// of one strip, which between them carry all 16 sets.  Everything a row
// needs stays in registers: the last five dgd rows (and squares), the
// radius-1 and radius-2 box sums of the row they centre on (vertical sums
// in the lane, horizontal neighbours by DPP wave shifts), per set the last
// three rows of the radius-1 pass's A/B and the last two odd rows of the
// radius-2 pass's, and the set's five projection sums.  Output columns are
// lanes 3 .. 60 (every lane a filter reaches is then inside the wave), so
// strips advance by 58 columns and may cross unit boundaries: each lane
// knows its unit, and at the end of its rows the wave adds its lanes'
// sums per unit into `out` with 64-bit atomics (integer, so the order does
// not matter).  A zero pass before and a division pass after turn the sums
// into the sums / unit area the projection solver reads.
//
// The filter is evaluated through T = 256 - A (a 256-entry u8 table in LDS,
// indexed by z) and B' = B + 8: with weights w summing to 2^(sh - 4),
//   rpot(sum w A * d + sum w B, sh) - 16 d = (sum w B' - (sum w T) * d) >> sh
// exactly, since sum w A * d = 16 d * 2^sh - (sum w T) * d and the +8 per
// weight unit is the rounding term 2^(sh - 1).
//
// Bounds (bd <= 10, samples inside the bit depth): box sums S1 <= 9207,
// S2 <= 25575; the variance terms p1 <= 1.32e6, p2 <= 1.03e7 (< 2^24) and
// p * s + 2^19 < 2^32 (the reference's own u32 product), T * S < 2^23 and
// T * S * one_by_x + 34816 < 2^31; sum w T <= 8160, sum w B' < 2^23, so the
// filter is one i24 multiply-add; |f|, |sv| <= 16368, so a product is below
// 2^28 and a lane's i32 partial holds 8 rows at 10 bits (64 at 8 bits)
// before it is folded into the i64 sum.
constexpr int kSgHalo = 3, kSgStrip = 64 - 2 * kSgHalo;
constexpr int kSgBoth = 0, kSgR1 = 1, kSgR0 = 2;

namespace {
// lane i receives lane i - 1 / i + 1 of the wave (DPP wave_shr:1 / wave_shl:1;
// the end lanes receive 0 and are never output lanes)
__device__ __forceinline__ int32_t lane_prev(int32_t v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t lane_next(int32_t v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t sides(int32_t v) { return lane_prev(v) + lane_next(v); }
__device__ __forceinline__ uint32_t mad_u24(uint32_t x, uint32_t y, uint32_t a) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "v"(a));
  return r;
}

template <int KIND>
struct SgSet {
  static constexpr bool R1 = KIND != kSgR0, R0 = KIND != kSgR1;
  uint32_t s0 = 0, s1 = 0;                                    // scales of the two passes
  int32_t t1a = 0, t1b = 0, t1c = 0, b1a = 0, b1b = 0, b1c = 0;  // radius 1: rows m-2, m-1, m
  int32_t t0a = 0, t0b = 0, b0a = 0, b0b = 0;                 // radius 2: older / newer odd row
  int32_t p00 = 0, p01 = 0, p11 = 0, pc0 = 0, pc1 = 0;        // i32 partials
  int64_t a00 = 0, a01 = 0, a11 = 0, ac0 = 0, ac1 = 0;

  __device__ __forceinline__ void fold() {
    if constexpr (R0) { a00 += (uint32_t)p00; ac0 += pc0; p00 = pc0 = 0; }
    if constexpr (R1) { a11 += (uint32_t)p11; ac1 += pc1; p11 = pc1 = 0; }
    if constexpr (R0 && R1) { a01 += p01; p01 = 0; }
  }
  // One row step centred on row m (ODD: m is an odd row of the unit, where
  // the radius-2 pass has A/B); with OUT, output row m - 1 (sample d, source
  // term sv, nd = -d).
  template <bool ODD, bool OUT>
  __device__ __forceinline__ void step(const uint8_t* tab, uint32_t p1, uint32_t S1, uint32_t p2, uint32_t S2,
                                       int32_t nd, int32_t sv) {
    if constexpr (R1) {
      const uint32_t z = min(mad_u24(p1, s1, 1u << 19) >> 20, 255u);
      const int32_t T = tab[z];
      const int32_t B = (int32_t)(mad_u24(mul_u24((uint32_t)T, S1), 455u, 34816u) >> 12);
      t1a = t1b; t1b = t1c; t1c = T;
      b1a = b1b; b1b = b1c; b1c = B;
    }
    if constexpr (R0 && ODD) {
      const uint32_t z = min(mad_u24(p2, s0, 1u << 19) >> 20, 255u);
      const int32_t T = tab[z];
      const int32_t B = (int32_t)(mad_u24(mul_u24((uint32_t)T, S2), 164u, 34816u) >> 12);
      t0a = t0b; t0b = T;
      b0a = b0b; b0b = B;
    }
    if constexpr (OUT) {
      int32_t f0 = 0, f1 = 0;
      if constexpr (R1) {  // weights 4 (centre, cross) and 3 (corners), sum 32
        const int32_t vT = t1a + t1b + t1c, vB = b1a + b1b + b1c;
        const int32_t wT = 4 * vT + 3 * sides(vT) + sides(t1b);
        const int32_t wB = 4 * vB + 3 * sides(vB) + sides(b1b);
        f1 = mad_i24(wT, nd, wB) >> 9;
      }
      if constexpr (R0) {
        if constexpr (ODD) {  // even output row: the odd rows above and below, 6 / 5, sum 32
          const int32_t vT = t0a + t0b, vB = b0a + b0b;
          f0 = mad_i24(6 * vT + 5 * sides(vT), nd, 6 * vB + 5 * sides(vB)) >> 9;
        } else {  // odd output row: its own row, 6 / 5, sum 16
          f0 = mad_i24(6 * t0b + 5 * sides(t0b), nd, 6 * b0b + 5 * sides(b0b)) >> 8;
        }
      }
      if constexpr (R0) { p00 = mad_i24(f0, f0, p00); pc0 = mad_i24(f0, sv, pc0); }
      if constexpr (R1) { p11 = mad_i24(f1, f1, p11); pc1 = mad_i24(f1, sv, pc1); }
      if constexpr (R0 && R1) p01 = mad_i24(f0, f1, p01);
    }
  }
};

// Wave sum of a signed 64-bit value whose lanes are below 2^40 in magnitude:
// the low 24 bits and the rest are summed as 32-bit values over each 16-lane
// row (DPP) and the four row results read into scalars.
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
  const int32_t lo = dpp_sum_i32<16>((int32_t)((uint32_t)v & 0xffffffu));
  const int32_t hi = dpp_sum_i32<16>((int32_t)(v >> 24));
  const int64_t l = (int64_t)__builtin_amdgcn_readlane(lo, 0) + __builtin_amdgcn_readlane(lo, 16) +
                    __builtin_amdgcn_readlane(lo, 32) + __builtin_amdgcn_readlane(lo, 48);
  const int64_t h = (int64_t)__builtin_amdgcn_readlane(hi, 0) + __builtin_amdgcn_readlane(hi, 16) +
                    __builtin_amdgcn_readlane(hi, 32) + __builtin_amdgcn_readlane(hi, 48);
  return h * (1 << 24) + l;
}
}  // namespace

template <typename Pix, int KA, int KB, int KC, int KD>
__device__ __forceinline__ void sgr_strip_wave(const Pix* __restrict__ dgd, const Pix* __restrict__ src, int pw,
                                               int ph, int x0, int vs, int ve, int uy, int nh, int unit_size,
                                               const uint8_t* tab, int4 sets, int64_t* __restrict__ out) {
  constexpr int BD = sizeof(Pix) == 1 ? 8 : 10, CS = BD - 8;
  constexpr int FOLD = BD == 8 ? 32 : 4;  // row pairs per i32 partial
  const int lane = threadIdx.x & 63;
  const int col = x0 - kSgHalo + lane, gx = clampi(col, 0, pw - 1);
  SgSet<KA> sa;
  SgSet<KB> sb;
  SgSet<KC> sc;
  SgSet<KD> sd;
  sa.s0 = kSgrParams[sets.x][1], sa.s1 = kSgrParams[sets.x][3];
  sb.s0 = kSgrParams[sets.y][1], sb.s1 = kSgrParams[sets.y][3];
  sc.s0 = kSgrParams[sets.z][1], sc.s1 = kSgrParams[sets.z][3];
  sd.s0 = kSgrParams[sets.w][1], sd.s1 = kSgrParams[sets.w][3];
  auto dg = [&](int y) -> int32_t { return dgd[(long)clampi(y, 0, ph - 1) * pw + gx]; };
  auto sr = [&](int y) -> int32_t { return src[(long)clampi(y, 0, ph - 1) * pw + gx]; };
  // dgd rows m-2 .. m+2 of the current step (d0..d4) and their squares
  int32_t d0 = 0, d1 = dg(vs - 3), d2 = dg(vs - 2), d3 = dg(vs - 1), d4 = dg(vs);
  uint32_t q0 = 0, q1 = mul_u24(d1, d1), q2 = mul_u24(d2, d2), q3 = mul_u24(d3, d3), q4 = mul_u24(d4, d4);
  auto step = [&](auto odd_c, auto out_c, int32_t nrow, int32_t srow) {
    constexpr bool ODD = decltype(odd_c)::value, OUT = decltype(out_c)::value;
    d0 = d1; d1 = d2; d2 = d3; d3 = d4; d4 = nrow;
    q0 = q1; q1 = q2; q2 = q3; q3 = q4; q4 = mul_u24(nrow, nrow);
    const int32_t c3s = d1 + d2 + d3, c3q = (int32_t)(q1 + q2 + q3);
    const int32_t S1 = c3s + sides(c3s), Q1 = c3q + sides(c3q);
    const uint32_t a1 = CS ? (uint32_t)(Q1 + (1 << (2 * CS - 1))) >> (2 * CS) : (uint32_t)Q1;
    const uint32_t e1 = CS ? (uint32_t)(S1 + (1 << (CS - 1))) >> CS : (uint32_t)S1;
    const uint32_t p1 = (uint32_t)max((int32_t)mad_u24(a1, 9u, 0u - mul_u24(e1, e1)), 0);
    uint32_t p2 = 0, S2 = 0;
    if constexpr (ODD) {
      const int32_t c5s = c3s + d0 + d4, c5q = c3q + (int32_t)(q0 + q4);
      const int32_t s5 = sides(c5s + sides(c5s)) - c5s, q5 = sides(c5q + sides(c5q)) - c5q;
      const uint32_t a2 = CS ? (uint32_t)(q5 + (1 << (2 * CS - 1))) >> (2 * CS) : (uint32_t)q5;
      const uint32_t e2 = CS ? (uint32_t)(s5 + (1 << (CS - 1))) >> CS : (uint32_t)s5;
      p2 = (uint32_t)max((int32_t)mad_u24(a2, 25u, 0u - mul_u24(e2, e2)), 0);
      S2 = (uint32_t)s5;
    }
    int32_t nd = 0, sv = 0;
    if constexpr (OUT) {
      nd = -d1;
      sv = (srow - d1) * 16;
    }
    sa.template step<ODD, OUT>(tab, p1, (uint32_t)S1, p2, S2, nd, sv);
    sb.template step<ODD, OUT>(tab, p1, (uint32_t)S1, p2, S2, nd, sv);
    sc.template step<ODD, OUT>(tab, p1, (uint32_t)S1, p2, S2, nd, sv);
    sd.template step<ODD, OUT>(tab, p1, (uint32_t)S1, p2, S2, nd, sv);
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  // pair 0: centres vs - 1 (odd) and vs, no output yet
  int32_t n0 = dg(vs + 1), n1 = dg(vs + 2);
  int32_t m0 = dg(vs + 3), m1 = dg(vs + 4), r0 = sr(vs), r1 = sr(vs + 1);
  step(T_{}, F_{}, n0, 0);
  step(F_{}, F_{}, n1, 0);
  int cnt = 0;
  // pair p >= 1: centres m = vs - 1 + 2p (odd) and m + 1, output rows m - 1, m
  for (int m = vs + 1; m <= ve; m += 2) {
    n0 = m0, n1 = m1;
    const int32_t s0 = r0, s1 = r1;
    m0 = dg(m + 4), m1 = dg(m + 5), r0 = sr(m + 1), r1 = sr(m + 2);  // next pair
    step(T_{}, T_{}, n0, s0);
    if (m + 1 > ve) break;
    step(F_{}, T_{}, n1, s1);
    if (++cnt == FOLD) {
      cnt = 0;
      sa.fold(), sb.fold(), sc.fold(), sd.fold();
    }
  }
  sa.fold(), sb.fold(), sc.fold(), sd.fold();
  // per unit the strip's output columns reach: wave sums into `out`
  const bool on = lane >= kSgHalo && lane < 64 - kSgHalo && col < pw;
  const int last = min(x0 + kSgStrip, pw) - 1;
  const int ux_lane = on ? min(col / unit_size, nh - 1) : -1;
  const int ux_hi = min(last / unit_size, nh - 1);
  for (int ux = min(x0 / unit_size, nh - 1); ux <= ux_hi; ux++) {
    const bool mine = ux_lane == ux;
    int64_t* o = out + (long)(uy * nh + ux) * 16 * 5;
    auto put = [&](int set, int m, int64_t v) {
      const int64_t t = wave_sum_i64(mine ? v : 0);
      if (lane == 0) atomicAdd((unsigned long long*)&o[set * 5 + m], (unsigned long long)t);
    };
    auto emit = [&](auto& s, int set) {
      using S = std::remove_reference_t<decltype(s)>;
      if constexpr (S::R0) { put(set, 0, s.a00); put(set, 3, s.ac0); }
      if constexpr (S::R1) { put(set, 2, s.a11); put(set, 4, s.ac1); }
      if constexpr (S::R0 && S::R1) put(set, 1, s.a01);
    };
    emit(sa, sets.x);
    emit(sb, sets.y);
    emit(sc, sets.z);
    emit(sd, sets.w);
  }
}

// Grid: one workgroup of four waves per (unit row half, 58-column strip); the
// waves carry sets {0,1,2,14}, {3,4,5,15}, {6,7,10,11}, {8,9,12,13}, which
// evens out their work (three two-pass sets and one radius-2 set against two
// two-pass and two radius-1 sets).
template <typename Pix>
__global__ __launch_bounds__(256) void k_sgr_strip(const Pix* __restrict__ dgd, const Pix* __restrict__ src,
                                                   int pw, int ph, int unit_size, int ss_y, int nh, int nstrips,
                                                   int64_t* __restrict__ out) {
  __shared__ uint8_t tab[256];  // 256 - x_by_xplus1(z)
  tab[threadIdx.x] = (uint8_t)(256 - x_by_xplus1(threadIdx.x));
  __syncthreads();
  // each unit row in two chunks of rows (the first an even number of rows
  // long, so the radius-2 pass keeps its odd rows): twice the waves, a
  // shorter tail
  const int b = xcd_chunked(blockIdx.x, gridDim.x);
  const int uy = b / (2 * nstrips), rem = b - uy * 2 * nstrips, half = rem / nstrips, sx = rem - half * nstrips;
  int hs, he, vs0, ve0;
  unit_limits(unit_size, pw, ph, ss_y, 0, uy, hs, he, vs0, ve0);
  const int mid = vs0 + (((ve0 - vs0) >> 1) & ~1);
  const int vs = half ? mid : vs0, ve = half ? ve0 : mid;
  if (vs >= ve) return;
  const int x0 = sx * kSgStrip, w = threadIdx.x >> 6;
  if (w == 0)
    sgr_strip_wave<Pix, kSgBoth, kSgBoth, kSgBoth, kSgR0>(dgd, src, pw, ph, x0, vs, ve, uy, nh, unit_size, tab,
                                                          make_int4(0, 1, 2, 14), out);
  else if (w == 1)
    sgr_strip_wave<Pix, kSgBoth, kSgBoth, kSgBoth, kSgR0>(dgd, src, pw, ph, x0, vs, ve, uy, nh, unit_size, tab,
                                                          make_int4(3, 4, 5, 15), out);
  else if (w == 2)
    sgr_strip_wave<Pix, kSgBoth, kSgBoth, kSgR1, kSgR1>(dgd, src, pw, ph, x0, vs, ve, uy, nh, unit_size, tab,
                                                        make_int4(6, 7, 10, 11), out);
  else
    sgr_strip_wave<Pix, kSgBoth, kSgBoth, kSgR1, kSgR1>(dgd, src, pw, ph, x0, vs, ve, uy, nh, unit_size, tab,
                                                        make_int4(8, 9, 12, 13), out);
}

__global__ __launch_bounds__(256) void k_fill_i64(int64_t* p, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += gridDim.x * 256L) p[i] = 0;
}

// Sums / unit area (C division, toward zero); an empty unit reads 0.
__global__ __launch_bounds__(256) void k_sgr_finalize(int64_t* out, int pw, int ph, int unit_size, int ss_y, int nh,
                                                      int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int u = i / 80, uy = u / nh, ux = u - uy * nh;
  int hs, he, vs, ve;
  unit_limits(unit_size, pw, ph, ss_y, ux, uy, hs, he, vs, ve);
  const int64_t size = (int64_t)(he - hs) * (ve - vs);
  out[i] = size > 0 ? out[i] / size : 0;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t deblock_plane(void* buf, int stride, int pw, int ph, int bd, int plane, int ss, int level_v,
                         int level_h, int sharpness, const DeblockGeomDev& g, hipStream_t st) {
  const Thr tv = make_thr(level_v, sharpness, bd), th = make_thr(level_h, sharpness, bd);
  const long nv = (long)(pw >> 2) * (ph & ~3), nhz = (long)(ph >> 2) * (pw & ~3);
  if (level_v && nv > 0) {
    const unsigned gr = (unsigned)((nv + 255) / 256);
    if (bd == 8)
      hipLaunchKernelGGL((k_deblock<uint8_t, true>), dim3(gr), dim3(256), 0, st, (uint8_t*)buf, stride, pw, ph,
                         plane, ss, level_v, tv, g);
    else
      hipLaunchKernelGGL((k_deblock<uint16_t, true>), dim3(gr), dim3(256), 0, st, (uint16_t*)buf, stride, pw, ph,
                         plane, ss, level_v, tv, g);
  }
  if (level_h && nhz > 0) {
    const unsigned gr = (unsigned)((nhz + 255) / 256);
    if (bd == 8)
      hipLaunchKernelGGL((k_deblock<uint8_t, false>), dim3(gr), dim3(256), 0, st, (uint8_t*)buf, stride, pw, ph,
                         plane, ss, level_h, th, g);
    else
      hipLaunchKernelGGL((k_deblock<uint16_t, false>), dim3(gr), dim3(256), 0, st, (uint16_t*)buf, stride, pw, ph,
                         plane, ss, level_h, th, g);
  }
  return hipGetLastError();
}

template <typename Pix>
static void trials_t(const Pix* recon, const Pix* src, Pix* work, int pw, int ph, int plane, int ss,
                     const LevelList& levels, int n_levels, int sharpness, int bd, const DeblockGeomDev& g,
                     uint64_t* sse, hipStream_t st) {
  const long n = (long)pw * ph;
  const unsigned cg = (unsigned)std::min<long>((n + 255) / 256, 2048);
  hipLaunchKernelGGL((k_trial_copy<Pix>), dim3(cg, n_levels), dim3(256), 0, st, recon, work, n);
  const long nv = (long)(pw >> 2) * (ph & ~3), nhz = (long)(ph >> 2) * (pw & ~3);
  if (nv > 0)
    hipLaunchKernelGGL((k_deblock_trials<Pix, true>), dim3((unsigned)((nv + 255) / 256), n_levels), dim3(256), 0, st,
                       work, pw, ph, plane, ss, levels, sharpness, bd, g);
  if (nhz > 0)
    hipLaunchKernelGGL((k_deblock_trials<Pix, false>), dim3((unsigned)((nhz + 255) / 256), n_levels), dim3(256), 0,
                       st, work, pw, ph, plane, ss, levels, sharpness, bd, g);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, sse, n_levels);
  hipLaunchKernelGGL((k_trial_sse<Pix>), dim3(cg, n_levels), dim3(256), 0, st, src, work, n,
                     (unsigned long long*)sse);
}

hipError_t deblock_trials(const void* recon, const void* src, void* work, int pw, int ph, int bd, int plane, int ss,
                          const uint8_t* levels_host, int n_levels, int sharpness, const DeblockGeomDev& g,
                          uint64_t* sse, hipStream_t st) {
  if (n_levels <= 0 || n_levels > 64) return hipErrorInvalidValue;
  LevelList levels{};
  for (int i = 0; i < n_levels; i++) levels.l[i] = levels_host[i];
  if (bd == 8)
    trials_t<uint8_t>((const uint8_t*)recon, (const uint8_t*)src, (uint8_t*)work, pw, ph, plane, ss, levels,
                      n_levels, sharpness, bd, g, sse, st);
  else
    trials_t<uint16_t>((const uint16_t*)recon, (const uint16_t*)src, (uint16_t*)work, pw, ph, plane, ss, levels,
                       n_levels, sharpness, bd, g, sse, st);
  return hipGetLastError();
}

hipError_t cdef_search(const void* rec_y, const void* rec_u, const void* rec_v, const void* src_y,
                       const void* src_u, const void* src_v, int width, int height, int bd, int damping, int sub_y,
                       int sub_uv, const uint8_t* skip, int mi_cols, int mi_rows, int32_t* dirvar, uint64_t* mse_y,
                       uint64_t* mse_uv, hipStream_t st) {
  const int nhfb = (width + 63) / 64, nvfb = (height + 63) / 64;
  const unsigned gr = (unsigned)(nhfb * nvfb);
  if (bd == 8) {
    hipLaunchKernelGGL((k_cdef_luma<uint8_t>), dim3(gr), dim3(256), 0, st, (const uint8_t*)rec_y,
                       (const uint8_t*)src_y, width, height, bd, damping, sub_y, nhfb, skip, mi_cols, mi_rows, mse_y, dirvar);
    if (mse_uv)
      hipLaunchKernelGGL((k_cdef_chroma<uint8_t>), dim3(gr), dim3(256), 0, st, (const uint8_t*)rec_u,
                         (const uint8_t*)rec_v, (const uint8_t*)src_u, (const uint8_t*)src_v, width, height, bd,
                         damping, sub_uv, nhfb, skip, mi_cols, mi_rows, (const int32_t*)dirvar, mse_uv);
  } else {
    hipLaunchKernelGGL((k_cdef_luma<uint16_t>), dim3(gr), dim3(256), 0, st, (const uint16_t*)rec_y,
                       (const uint16_t*)src_y, width, height, bd, damping, sub_y, nhfb, skip, mi_cols, mi_rows, mse_y, dirvar);
    if (mse_uv)
      hipLaunchKernelGGL((k_cdef_chroma<uint16_t>), dim3(gr), dim3(256), 0, st, (const uint16_t*)rec_u,
                         (const uint16_t*)rec_v, (const uint16_t*)src_u, (const uint16_t*)src_v, width, height, bd,
                         damping, sub_uv, nhfb, skip, mi_cols, mi_rows, (const int32_t*)dirvar, mse_uv);
  }
  return hipGetLastError();
}

hipError_t wiener_stats(const void* dgd, const void* src, int pw, int ph, int bd, int win, int unit_size, int ss_y,
                        int nh, int nv, int64_t* M, int64_t* H, hipStream_t st) {
  const unsigned gr = (unsigned)(nh * nv);
#define ZW(P, WN)                                                                                              \
  hipLaunchKernelGGL((k_wiener_stats<P, WN>), dim3(gr), dim3(256), 0, st, (const P*)dgd, (const P*)src, pw, ph, \
                     unit_size, ss_y, nh, make_int4(0, 0, 0, 0), M, H)
  if (bd == 8) {
    if (win == 7) ZW(uint8_t, 7);
    else ZW(uint8_t, 5);
  } else {
    if (win == 7) ZW(uint16_t, 7);
    else ZW(uint16_t, 5);
  }
#undef ZW
  return hipGetLastError();
}

hipError_t unit_sse(const void* a, const void* b, int pw, int ph, int bd, int unit_size, int ss_y, int nh, int nv,
                    uint64_t* out, hipStream_t st) {
  const unsigned gr = (unsigned)(nh * nv);
  if (bd == 8)
    hipLaunchKernelGGL((k_unit_sse<uint8_t>), dim3(gr), dim3(256), 0, st, (const uint8_t*)a, (const uint8_t*)b, pw,
                       ph, unit_size, ss_y, nh, out);
  else
    hipLaunchKernelGGL((k_unit_sse<uint16_t>), dim3(gr), dim3(256), 0, st, (const uint16_t*)a, (const uint16_t*)b,
                       pw, ph, unit_size, ss_y, nh, out);
  return hipGetLastError();
}

hipError_t sgr_stats(const void* dgd, const void* src, int pw, int ph, int bd, int unit_size, int ss_y, int nh,
                     int nv, int64_t* out, hipStream_t st) {
  const int nstrips = (pw + kSgStrip - 1) / kSgStrip, n = nh * nv * 80;
  hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)std::min((n + 255) / 256, 1024)), dim3(256), 0, st, out, (long)n);
  const unsigned gr = (unsigned)(nv * nstrips * 2);
  if (bd == 8)
    hipLaunchKernelGGL((k_sgr_strip<uint8_t>), dim3(gr), dim3(256), 0, st, (const uint8_t*)dgd, (const uint8_t*)src,
                       pw, ph, unit_size, ss_y, nh, nstrips, out);
  else
    hipLaunchKernelGGL((k_sgr_strip<uint16_t>), dim3(gr), dim3(256), 0, st, (const uint16_t*)dgd,
                       (const uint16_t*)src, pw, ph, unit_size, ss_y, nh, nstrips, out);
  hipLaunchKernelGGL(k_sgr_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, pw, ph, unit_size,
                     ss_y, nh, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain form: one call of the reference's per-block kernels on staged data.
// ---------------------------------------------------------------------------
// lpf_{horizontal,vertical}_{4,6,8,14}(_hbd): the 4 lines of the call packed
// line-major (len samples each, p side first).
template <typename Pix>
__global__ __launch_bounds__(64) void k_plain_lpf(Pix* __restrict__ seg, int len, Thr t) {
  const int i = threadIdx.x;
  if (i >= 4) return;
  if (len == 4) filter_at<Pix, 4>(seg, (long)i * len, 1, t);
  else if (len == 6) filter_at<Pix, 6>(seg, (long)i * len, 1, t);
  else if (len == 8) filter_at<Pix, 8>(seg, (long)i * len, 1, t);
  else filter_at<Pix, 14>(seg, (long)i * len, 1, t);
}

// cdef_find_dir on one packed 8x8: out = {dir, var}.
__global__ __launch_bounds__(64) void k_plain_cdef_dir(const uint16_t* __restrict__ blk, int cs,
                                                       int32_t* __restrict__ out) {
  if (threadIdx.x != 0) return;
  int var = 0;
  out[0] = find_dir<8>(blk, cs, &var);
  out[1] = var;
}

// cdef_filter_block(_hbd): tile = the call's padded input, rows -kVB .. and
// columns -kHB .. around the block, row stride kTS; one thread per output
// sample of the rows the row step keeps.
__global__ __launch_bounds__(64) void k_plain_cdef_filter(const uint16_t* __restrict__ tile, PlainCdef p,
                                                          int32_t* __restrict__ out) {
  const int t = threadIdx.x;
  if (t >= p.rows * p.cols) return;
  const int i = t / p.cols, j = t - (t / p.cols) * p.cols;
  if (i % p.sub) return;
  const uint16_t* cp = tile + (i + kVB) * kTS + j + kHB;
  PixTaps pt;
  gather_taps(cp, p.dir, pt);
  int32_t sum = p.pri ? pri_sum(pt, p.pri, (p.pri >> p.cs) & 1, p.pdamp) : 0;
  if (p.sec) {
    const int32_t sh = max(0, p.sdamp - msb32((uint32_t)p.sec));
#pragma unroll
    for (int m = 0; m < 8; m++) sum += (m < 4 ? 2 : 1) * constrain(pt.d_s[m], p.sec, sh);
  }
  out[t] = cdef_out((int16_t)cp[0], sum, pt.mn, pt.mx);
}

hipError_t plain_lpf(void* seg, int hbd, int len, int mblim, int lim, int hev, int bd, hipStream_t st) {
  Thr t;
  t.mblim = mblim;
  t.lim = lim;
  t.hev = hev;
  t.shift = bd - 8;
  t.bd = bd;
  if (hbd)
    hipLaunchKernelGGL((k_plain_lpf<uint16_t>), dim3(1), dim3(64), 0, st, (uint16_t*)seg, len, t);
  else
    hipLaunchKernelGGL((k_plain_lpf<uint8_t>), dim3(1), dim3(64), 0, st, (uint8_t*)seg, len, t);
  return hipGetLastError();
}

hipError_t plain_cdef_dir(const uint16_t* blk, int cs, int32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_plain_cdef_dir, dim3(1), dim3(64), 0, st, blk, cs, out);
  return hipGetLastError();
}

hipError_t plain_cdef_filter(const uint16_t* tile, const PlainCdef& p, int32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_plain_cdef_filter, dim3(1), dim3(64), 0, st, tile, p, out);
  return hipGetLastError();
}

hipError_t plain_wiener_window(const void* dgd, const void* src, int pw, int ph, int bd, int win, int hs, int he,
                               int vs, int ve, int64_t* M, int64_t* H, hipStream_t st) {
  const int4 lim = make_int4(hs, he, vs, ve);
#define ZP(P, WN)                                                                                              \
  hipLaunchKernelGGL((k_wiener_stats<P, WN>), dim3(1), dim3(256), 0, st, (const P*)dgd, (const P*)src, pw, ph, \
                     64, 0, 1, lim, M, H)
  if (bd == 8) {
    if (win == 7) ZP(uint8_t, 7);
    else ZP(uint8_t, 5);
  } else {
    if (win == 7) ZP(uint16_t, 7);
    else ZP(uint16_t, 5);
  }
#undef ZP
  return hipGetLastError();
}

}  // namespace zav1
