// Partition pre-pass (light PD0) of the allintra presets for gfx950: every
// 64x64 unit of a batch of stills gets its variance map, its pass level, the
// cost of every square block the pass may test and the parent-first tree
// decision with its node records.
//
// Reference: rust/crates/svtav1-encoder/src/pd0.rs -- compute_b64_variance
// :81, qp_th_scaling_factors :185, max_block_size_allintra :355,
// pd0_detector_allintra_demotes :441, lvl6_cost_allintra :456, tx_quant_core
// :697, cost_coeffs_txb_pd0 :1019, check_is_subres_safe :1075, the LVL_0 /
// LVL_1 / LVL_5 block costs :1494-1690 and Pd0Ctx::pick :1702 behind the
// entry points :1890 (eff-M9), :1990 (LVL_0) and :2144 (LVL_1).
//
// The reference walks a unit depth first and costs a block when the walk
// reaches it.  A block's cost depends only on the source (DC prediction
// from source neighbours, the unit's level and subres decision), never on
// the walk, so here every block the walk may reach is costed in one launch
// per block size over all units, and the walk itself -- early exits, the
// parent-versus-split compare -- runs afterwards over the stored costs,
// one wave per unit.  The subres decision (LVL_5) is the 64x64 block's and
// is taken before any cost, as the walk takes it first.
//
// Scope: units are complete 64x64 blocks (image sides multiples of 64), so
// the specification's forced edge splits never arise.  Samples are 8-bit:
// a 10-bit plane enters MSB-truncated (v >> 2), as the reference's pass
// reads it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "coeff_ctx.h"
#include "zav1_dev.h"
#include "zav1_kernels.h"

namespace zav1 {
namespace {
using namespace cctx;

constexpr int kLog2Lut[5] = {6, 5, 4, 3, 2};
constexpr int kGridLut[5] = {1, 2, 4, 8, 16};
constexpr int kBaseLut[5] = {0, 1, 5, 21, 85};

__host__ __device__ constexpr int lg2(int v) { return v <= 1 ? 0 : 1 + lg2(v >> 1); }
__host__ __device__ inline int node_index(int sq, int ox, int oy) {
  const int l = 6 - lg2(sq);
  return kBaseLut[l] + (oy >> kLog2Lut[l]) * kGridLut[l] + (ox >> kLog2Lut[l]);
}
__host__ __device__ inline uint64_t div_round(uint64_t x, uint64_t y) { return (x + (y >> 1)) / y; }
__host__ __device__ inline uint64_t rdc(uint64_t lambda, uint64_t rate, uint64_t dist) {
  return ((rate * lambda + 256) >> 9) + (dist << 7);
}
// TxSize of a w x h transform
__host__ __device__ constexpr int tx_of(int w, int h) {
  for (int t = 0; t < 19; t++)
    if (tx_w(t) == w && tx_h(t) == h) return t;
  return -1;
}

template <typename Pix>
__device__ __forceinline__ int s8(const Pix* p, long i, int sh) {
  return (int)p[i] >> sh;
}

// ---- per unit: variance map, level, depth cap, subres, LVL_6 costs ----------
template <typename Pix>
__global__ __launch_bounds__(256) void k_pd0_prep(const Pix* __restrict__ plane, Pd0Params p,
                                                  Pd0SbInfo* __restrict__ info, Pd0Node* __restrict__ nodes) {
  const int s = blockIdx.x, tid = threadIdx.x, sbw = p.width >> 6;
  const int x0 = (s % sbw) << 6, y0 = (s / sbw) << 6;  // stacked plane
  const int sh = p.bd - 8;
  const long W = p.width;
  __shared__ uint32_t s_sum[64], s_sq[64];
  __shared__ uint16_t s_var[85];
  __shared__ uint32_t s_red[4];
  __shared__ int s_mode, s_max, s_dc;
  if (tid < 64) s_sum[tid] = s_sq[tid] = 0;
  if (tid < 4) s_red[tid] = 0;
  __syncthreads();
  {
    // even rows of each 8x8: thread = (even row k, 8-column chunk)
    const int row = 2 * (tid >> 3), ch = tid & 7;
    const Pix* r = plane + (long)(y0 + row) * W + x0 + ch * 8;
    uint32_t sum = 0, sq = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t v = (uint32_t)r[c] >> sh;
      sum += v;
      sq += v * v;
    }
    const int b = (row >> 3) * 8 + ch;
    atomicAdd(&s_sum[b], sum);
    atomicAdd(&s_sq[b], sq);
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t m8[64], q8[64], m16[16], q16[16], m32[4], q32[4];
    for (int i = 0; i < 64; i++) {
      m8[i] = (uint64_t)s_sum[i] << 3;
      q8[i] = (uint64_t)s_sq[i] << 11;
    }
    for (int by = 0; by < 4; by++)
      for (int bx = 0; bx < 4; bx++) {
        const int f = by * 16 + bx * 2;
        m16[by * 4 + bx] = (m8[f] + m8[f + 1] + m8[f + 8] + m8[f + 9]) >> 2;
        q16[by * 4 + bx] = (q8[f] + q8[f + 1] + q8[f + 8] + q8[f + 9]) >> 2;
      }
    for (int by = 0; by < 2; by++)
      for (int bx = 0; bx < 2; bx++) {
        const int f = by * 8 + bx * 2;
        m32[by * 2 + bx] = (m16[f] + m16[f + 1] + m16[f + 4] + m16[f + 5]) >> 2;
        q32[by * 2 + bx] = (q16[f] + q16[f + 1] + q16[f + 4] + q16[f + 5]) >> 2;
      }
    const uint64_t m64 = (m32[0] + m32[1] + m32[2] + m32[3]) >> 2, q64 = (q32[0] + q32[1] + q32[2] + q32[3]) >> 2;
    s_var[0] = (uint16_t)((q64 - m64 * m64) >> 16);
    for (int i = 0; i < 4; i++) s_var[1 + i] = (uint16_t)((q32[i] - m32[i] * m32[i]) >> 16);
    for (int i = 0; i < 16; i++) s_var[5 + i] = (uint16_t)((q16[i] - m16[i] * m16[i]) >> 16);
    for (int i = 0; i < 64; i++) s_var[21 + i] = (uint16_t)((q8[i] - m8[i] * m8[i]) >> 16);
    // depth cap (pd0.rs:355) and level (pd0.rs:441 detector)
    const uint16_t cap = (uint16_t)div_round(7500ull * p.qw, p.qwd);
    const int capped = s_var[0] <= cap ? 64 : 32;
    int mode, mx;
    if (p.level == 0) {
      mode = 0;
      mx = capped;
    } else if (p.level == 1) {
      mode = 1;
      mx = p.cap_max_block ? capped : 64;
    } else {
      const int32_t v64 = s_var[0], v32 = (((int32_t)s_var[1] + s_var[2] + s_var[3] + s_var[4]) >> 2) * 4;
      int32_t a16 = 0;
      for (int i = 5; i < 21; i++) a16 += s_var[i];
      const int32_t v16 = (a16 >> 4) * 16, th = (int32_t)div_round(7500ull * p.qw, p.qwd);
      mode = (abs(v32 - v64) < th && abs(v16 - v32) < th) ? 5 : 6;
      mx = capped;
    }
    s_mode = mode;
    s_max = mx < p.max_tx ? mx : p.max_tx;
  }
  __syncthreads();
  const int mode = s_mode, max_sq = s_max;
  // LVL_5 subres: the 64x64 block's DC prediction from source neighbours,
  // SAD of even vs odd rows (pd0.rs:1075); only when the 64x64 is tested
  int subres = 0;
  if (mode == 5 && max_sq == 64) {
    const int yi = y0 % p.height;
    const bool ha = yi > 0, hl = x0 > 0;
    if (tid < 64) {
      uint32_t a = ha ? (uint32_t)s8(plane, (long)(y0 - 1) * W + x0 + tid, sh) : 0;
      uint32_t l = hl ? (uint32_t)s8(plane, (long)(y0 + tid) * W + x0 - 1, sh) : 0;
      atomicAdd(&s_red[0], a + l);
    }
    __syncthreads();
    if (tid == 0) {
      const uint32_t sum = s_red[0];
      s_dc = (ha && hl) ? (int)((sum + 64) >> 7) : ((ha || hl) ? (int)((sum + 32) >> 6) : 128);
    }
    __syncthreads();
    const int dc = s_dc;
    uint32_t ev = 0, od = 0;
    for (int i = tid; i < 4096; i += 256) {
      const int r = i >> 6, c = i & 63;
      const uint32_t d = (uint32_t)abs(s8(plane, (long)(y0 + r) * W + x0 + c, sh) - dc);
      if (r & 1) od += d;
      else ev += d;
    }
    atomicAdd(&s_red[1], ev);
    atomicAdd(&s_red[2], od);
    __syncthreads();
    const int64_t e1 = max((int64_t)s_red[1], (int64_t)1), o1 = max((int64_t)s_red[2], (int64_t)1);
    const int64_t dev = ((e1 - o1) * 100) / o1;
    subres = (dev < 0 ? -dev : dev) <= 5;
  }
  Pd0SbInfo* o = info + s;
  if (tid < 85) o->var[tid] = s_var[tid];
  if (tid == 0) {
    o->mode = (uint8_t)mode;
    o->max_sq = (uint8_t)max_sq;
    o->subres = (uint8_t)subres;
    o->pad[0] = o->pad[1] = o->pad[2] = 0;
  }
  // LVL_6 closed-form costs (pd0.rs:456) of the tested nodes
  if (mode == 6 && tid < 85) {
    const int l = tid < 1 ? 0 : (tid < 5 ? 1 : (tid < 21 ? 2 : 3));
    const int sq = 64 >> l, k = tid - kBaseLut[l], g = kGridLut[l];
    const int ox = (k % g) * sq, oy = (k / g) * sq;
    if (sq <= max_sq && sq >= p.min_sq) {
      const uint64_t qw = p.qw, qwd = p.qwd, bv = s_var[tid], area = (uint64_t)sq * sq;
      uint64_t bias = 1000;
      if (sq == 64) {
        bias += 50 * min(bv / div_round(100 * qw, qwd), (uint64_t)10);
      } else if (sq >= 16) {
        const int sl = l + 1, ss = kLog2Lut[sl], sb = kBaseLut[sl], sg = kGridLut[sl];
        const int i0 = sb + (oy >> ss) * sg + (ox >> ss);
        const uint64_t a = s_var[i0], b = s_var[i0 + 1], c = s_var[i0 + sg], d = s_var[i0 + sg + 1];
        const uint64_t mn = min(min(a, b), min(c, d)), mxv = max(max(a, b), max(c, d));
        bias += 25 * min(bv / div_round(400 * qw, qwd), (uint64_t)10);
        bias += 10 * min((mxv - mn) / div_round(25 * qw, qwd), (uint64_t)10);
      } else {
        bias += 40 * min(bv / div_round(25 * qw, qwd), (uint64_t)10);
      }
      nodes[(long)s * p.nps + tid].cost = area * bias / 1000;
    }
  }
}

// ---- block costs of one size over every unit ---------------------------------
// W x H transform: W = the block side, H = W >> step (LVL_5 subres).  A group
// of W lanes per block; the column pass runs on lanes < W, the row pass with
// its fold, quantize_b, eob and distortion on lanes < H, the LVL_1
// coefficient rate (the fast-level-2 loop) over the group in scan order.
template <int W, int H, typename Pix>
__global__ __launch_bounds__(256) void k_pd0_cost(const Pix* __restrict__ plane, Pd0Params p,
                                                  const Pd0SbInfo* __restrict__ info, const QuantRow* __restrict__ row_p,
                                                  const uint16_t* __restrict__ iscan_all,
                                                  const uint16_t* __restrict__ fscan_all,
                                                  const int32_t* __restrict__ costs, Pd0Node* __restrict__ nodes) {
  constexpr int STEP = W == H ? 0 : 1, T = W, JPW = 256 / T, LS = W + 1, NB = (64 / W) * (64 / W);
  constexpr int TXS = tx_of(W, H);
  constexpr int PW = W < 32 ? W : 32, PH = H < 32 ? H : 32, N = PW * PH;
  constexpr int CBC = fwd_cb_col(W, H), CBR = fwd_cb_row(W, H);
  constexpr FwdShift FS = fwd_shift(W, H);
  constexpr long long kBCol = 255LL << FS.s0;
  constexpr long long kBRow = fwd_row_bound(H, CBC, FS.s1, kBCol);
  constexpr int LSC = tx_log_scale(W, H);
  constexpr bool RECT2 = (W == 2 * H);
  __shared__ int32_t lds[JPW * H * LS];
  __shared__ int32_t tab[kTxbWords + kEobWords];
  __shared__ uint16_t sc[N];

  const int tid = threadIdx.x, jl = tid / T, t = tid - jl * T;
  const long n_jobs = (long)p.n_sb * NB;
  const long job = (long)blockIdx.x * JPW + jl;
  const int s = (int)((job < n_jobs ? job : 0) / NB), b = (int)((job < n_jobs ? job : 0) % NB);
  const Pd0SbInfo in = info[s];
  const int ox = (b % (64 / W)) * W, oy = (b / (64 / W)) * W;
  const int sbw = p.width >> 6;
  const int x = ((s % sbw) << 6) + ox, y = ((s / sbw) << 6) + oy;  // stacked plane
  const int mode = in.mode;
  const bool step_ok = mode == 5 ? (in.subres ? STEP == 1 : STEP == 0) : STEP == 0;
  const bool active = job < n_jobs && mode != 6 && step_ok && W <= in.max_sq && W >= p.min_sq;
  const bool any_rate = p.level == 1;
  if (any_rate) {
    const int32_t* tx_src = costs + c_txs_ctx(TXS) * 2 * kTxbWords;  // plane 0
    const int32_t* eob_src = costs + 10 * kTxbWords + c_eob_ms(TXS) * 2 * kEobWords;
    for (int i = tid; i < kTxbWords; i += 256) tab[i] = tx_src[i];
    for (int i = tid; i < kEobWords; i += 256) tab[kTxbWords + i] = eob_src[i];
    for (int i = tid; i < N; i += 256) sc[i] = fscan_all[TXS * 3072 + i];
  }
  const int sh = p.bd - 8;
  const long PWd = p.width;
  int32_t* blk = lds + jl * H * LS;

  // DC prediction from the source neighbours (tile-aware at LVL_1)
  int dc = 128;
  {
    const int yi = y % p.height;
    const int ty0 = (mode == 1 && p.tile_h > 0) ? yi - yi % p.tile_h : 0;
    const int tx0 = (mode == 1 && p.tile_w > 0) ? x - x % p.tile_w : 0;
    const bool ha = yi > ty0, hl = x > tx0;
    int32_t a = 0;
    if (active) {
      if (ha) a += s8(plane, (long)(y - 1) * PWd + x + t, sh);
      if (hl) a += s8(plane, (long)(y + t) * PWd + x - 1, sh);
    }
    a = group_sum_i32<T>(a);
    if (ha && hl) dc = (a + W) >> (lg2(W) + 1);
    else if (ha || hl) dc = (a + W / 2) >> lg2(W);
  }
  // column pass
  if (active) {
    int32_t v[H];
#pragma unroll
    for (int r = 0; r < H; r++) v[r] = (s8(plane, (long)(y + (r << STEP)) * PWd + x + t, sh) - dc) << FS.s0;
    fwd1d<H, CBC, kBCol>(v, 0);
#pragma unroll
    for (int r = 0; r < H; r++) blk[r * LS + t] = rshift_round32(v[r], FS.s1);
  }
  __syncthreads();
  // row pass, fold, quantize_b at the pass's quantizer, eob, distortion
  uint64_t fdist = 0;
  int32_t eob = 0;
  if (active && t < H) {
    int32_t v[W];
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = blk[t * LS + c];
    fwd1d<W, CBR, kBRow>(v, 0);
    const QuantRow& qr = *row_p;
    const uint16_t* iscan = iscan_all + (TXS * 3) * 1024 + (t < 32 ? t : 0) * PW;
#pragma unroll
    for (int c = 0; c < W; c++) {
      int32_t xv = rshift_round32(v[c], FS.s2);
      if constexpr (RECT2) xv = (__mul24(xv, 5793) + 2048) >> 12;
      if (t >= 32 || c >= 32) {
        fdist += (uint64_t)((int64_t)xv * xv);
        continue;
      }
      const int iz = (t | c) != 0;
      const int32_t sg = xv >> 31, a = (xv ^ sg) - sg;
      const int32_t q = quant_b_level(a, qr, iz, LSC, 8);
      int32_t dq = 0, qs = 0;
      if (q != 0) {
        dq = (dequant_level(q, qr, iz, LSC) ^ sg) - sg;
        qs = (q ^ sg) - sg;
        eob = max(eob, (int)iscan[c] + 1);
      }
      const int64_t e = (int64_t)xv - dq;
      fdist += (uint64_t)(e * e);
      blk[t * LS + c] = qs;  // the level map the rate reads (packed top-left)
    }
  }
  fdist = group_sum_u64<T>(fdist);
  eob = group_max_i32<T>(eob);
  constexpr int SHF = (1 - LSC) * 2;
  uint64_t dist = SHF < 0 ? (fdist << (-SHF)) : (fdist >> SHF);
  dist <<= STEP;
  uint64_t bits = 0;
  if (p.level == 1) {  // frame-uniform: every unit of a LVL_1 pass is LVL_1
    __syncthreads();
    constexpr int TH = (PW * PH) >> 5;
    if (p.coeff_rate_est_lvl >= 2 && eob < TH) {
      bits = 6000 + (uint64_t)eob * 500;
    } else if (eob == 0) {
      bits = (uint64_t)tab[kSkip + 1];
    } else {
      const int32_t* cost = tab;
      const int32_t* ecost = tab + kTxbWords;
      auto lv = [blk](int r, int c) -> int {
        if (r >= PH || c >= PW) return 0;
        return min(abs(blk[r * LS + c]), 127);
      };
      const int mid = min(eob - 2, eob / 2);
      int32_t sum = 0;
      if (active)
        for (int i = t; i < eob; i += T) {
          const bool is_eob = i == eob - 1;
          if (!is_eob && i != 0 && i > mid) continue;
          const int pos = sc[i], v = blk[(pos / PW) * LS + (pos % PW)], l = abs(v);
          const int ctx = nz_ctx<TXS>(lv, pos, i, is_eob, kCls2D);
          if (is_eob) sum += cost[kBaseEob + ctx * 3 + max(min(l, 3) - 1, 0)];
          else sum += cost[kBase + ctx * 8 + min(l, 3)];
          if (v != 0) {
            sum += pos == 0 ? cost[kDcSign + (v < 0)] : 512;
            if (l > 2) sum += cost[kLps + br_ctx_of<TXS>(lv, pos, kCls2D) * 26 + min(l - 3, 12)] + golomb_cost(l);
          }
        }
      sum = group_sum_i32<T>(sum);
      const int txt = W == 4 ? p.txt_dc[0] : (W == 8 ? p.txt_dc[1] : (W == 16 ? p.txt_dc[2] : 0));
      bits = (uint64_t)(sum + cost[kSkip] + txt + eob_cost_of(eob, cost, ecost, kCls2D));
    }
  } else {
    bits = 5000 + (uint64_t)p.ires * 1600 + 100 * (uint64_t)eob;
  }
  if (active && t == 0) {
    const uint64_t rate = bits + (uint64_t)p.skip0_bits + (uint64_t)p.none_bits;
    nodes[(long)s * p.nps + node_index(W, ox, oy)].cost = rdc(p.lambda, rate, dist);
  }
}

// ---- the walk (pd0.rs:1702) over the stored costs ------------------------------
struct PickCtx {
  const Pd0Node* nd;
  uint8_t* flags;   // LDS, per node
  uint8_t* leaf;    // LDS, 16 x 16 map of log2 leaf sizes
  uint64_t lambda;
  const int32_t* split_bits;
  int mode, max_sq, min_sq;
};
template <int SQ>
__device__ uint64_t pick(const PickCtx& c, int ox, int oy) {
  const int ni = node_index(SQ, ox, oy);
  uint8_t fl = kPd0Visited;
  const bool tested = SQ <= c.max_sq && SQ >= c.min_sq;
  const uint64_t pc = tested ? c.nd[ni].cost : 0;
  if (tested) fl |= kPd0Tested;
  bool leaf = true;
  uint64_t res = pc;
  if constexpr (SQ > 4) {
    if (SQ > c.min_sq) {
      uint64_t split = 0;
      if (c.mode == 1) split = rdc(c.lambda, (uint64_t)c.split_bits[lg2(SQ) - 3], 0);
      else if (c.mode != 6) split = rdc(c.lambda, 2 * (uint64_t)c.split_bits[lg2(SQ) - 3], 0);
      bool valid = true;
      for (int i = 0; i < 4; i++) {
        // early exits: pc * th * 1000 <= split * 10^6 (th 50, then 1000);
        // costs stay below 2^44, so the forms below do not overflow
        if (c.mode != 6 && tested && pc * (i == 0 ? 50u : 1000u) <= split * 1000u) {
          valid = false;
          break;
        }
        split += pick<SQ / 2>(c, ox + (i & 1) * (SQ / 2), oy + (i >> 1) * (SQ / 2));
      }
      if (valid && !(tested && pc <= split)) {
        fl |= kPd0Split;
        leaf = false;
        res = split;
      }
    }
  }
  c.flags[ni] = fl;
  if (leaf)
    for (int r = oy / 4; r < (oy + SQ) / 4; r++)
      for (int k = ox / 4; k < (ox + SQ) / 4; k++) c.leaf[r * 16 + k] = (uint8_t)lg2(SQ);
  return res;
}

__global__ __launch_bounds__(64) void k_pd0_pick(Pd0Params p, const Pd0SbInfo* __restrict__ info,
                                                 Pd0Node* __restrict__ nodes, uint8_t* __restrict__ leaf_map) {
  const int s = blockIdx.x, lane = threadIdx.x;
  __shared__ uint8_t fl[341];
  __shared__ uint8_t lm[256];
  __shared__ int32_t sb[4];
  for (int i = lane; i < p.nps; i += 64) fl[i] = 0;
  if (lane < 4) sb[lane] = p.split_bits[lane];
  __syncthreads();
  Pd0Node* nd = nodes + (long)s * p.nps;
  if (lane == 0) {
    const Pd0SbInfo in = info[s];
    PickCtx c{nd, fl, lm, p.lambda, sb, in.mode, in.max_sq, p.min_sq};
    pick<64>(c, 0, 0);
  }
  __syncthreads();
  for (int i = lane; i < p.nps; i += 64) {
    const uint8_t f = fl[i];
    Pd0Node o;
    o.cost = (f & kPd0Tested) ? nd[i].cost : 0;
    o.flags = f;
    for (int k = 0; k < 7; k++) o.pad[k] = 0;
    nd[i] = o;
  }
  for (int i = lane; i < 256; i += 64) leaf_map[(long)s * 256 + i] = lm[i];
}

template <int W, int H, typename Pix>
hipError_t launch_cost(const void* plane, const Pd0Params& p, const Pd0SbInfo* info, const QuantRow* row,
                       const uint16_t* iscan, const uint16_t* fscan, const int32_t* costs, Pd0Node* nodes,
                       hipStream_t st) {
  constexpr int JPW = 256 / W, NB = (64 / W) * (64 / W);
  const long n = (long)p.n_sb * NB;
  hipLaunchKernelGGL((k_pd0_cost<W, H, Pix>), dim3((unsigned)((n + JPW - 1) / JPW)), dim3(256), 0, st,
                     (const Pix*)plane, p, info, row, iscan, fscan, costs, nodes);
  return hipGetLastError();
}

template <typename Pix>
hipError_t run_pix(const void* plane, const Pd0Params& p, const QuantRow* row, const uint16_t* iscan,
                   const uint16_t* fscan, const int32_t* costs, Pd0SbInfo* info, Pd0Node* nodes, uint8_t* leaf_map,
                   hipStream_t st) {
  hipLaunchKernelGGL((k_pd0_prep<Pix>), dim3((unsigned)p.n_sb), dim3(256), 0, st, (const Pix*)plane, p, info, nodes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_cost<64, 64, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  if ((e = launch_cost<32, 32, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  if ((e = launch_cost<16, 16, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  if ((e = launch_cost<8, 8, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  if (p.min_sq <= 4)
    if ((e = launch_cost<4, 4, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  if (p.level == 5) {  // the subres forms of the units whose 64x64 passed the check
    if ((e = launch_cost<64, 32, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
    if ((e = launch_cost<32, 16, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
    if ((e = launch_cost<16, 8, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
    if ((e = launch_cost<8, 4, Pix>(plane, p, info, row, iscan, fscan, costs, nodes, st)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_pd0_pick, dim3((unsigned)p.n_sb), dim3(64), 0, st, p, info, nodes, leaf_map);
  return hipGetLastError();
}

}  // namespace

hipError_t pd0_run(const void* plane, const Pd0Params& p, const QuantRow* row, const uint16_t* iscan,
                   const uint16_t* fscan, const int32_t* costs, Pd0SbInfo* info, Pd0Node* nodes, uint8_t* leaf_map,
                   hipStream_t st) {
  if (p.n_sb <= 0) return hipSuccess;
  return p.bd == 8 ? run_pix<uint8_t>(plane, p, row, iscan, fscan, costs, info, nodes, leaf_map, st)
                   : run_pix<uint16_t>(plane, p, row, iscan, fscan, costs, info, nodes, leaf_map, st);
}

}  // namespace zav1
