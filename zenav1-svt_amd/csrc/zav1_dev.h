// Device-side shared definitions for the zav1 HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "txfm1d_gen.h"

namespace zav1 {

// Logical index of workgroup b of a grid of G with the logical range cut
// into 8 contiguous chunks, one per XCD: the dispatcher hands consecutive
// workgroups round-robin to the 8 XCDs (b runs on XCD b % 8), so chunking
// keeps neighbouring work items, which share halo and source lines, in one
// XCD's L2.  A bijection on [0, G).
__device__ __forceinline__ int xcd_chunked(int b, int G) {
  const int x = b & 7, i = b >> 3, per = G >> 3, rem = G & 7;
  return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// Orders one wave's LDS traffic between its lanes (a wave-level fence and
// barrier; lanes of other waves are not waited for).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// TxSize in C order (rust/crates/svtav1-types/src/transform.rs:7-29).
__host__ __device__ constexpr int tx_w(int s) {
  return s == 0 ? 4 : s == 1 ? 8 : s == 2 ? 16 : s == 3 ? 32 : s == 4 ? 64 : s == 5 ? 4 : s == 6 ? 8
       : s == 7 ? 8 : s == 8 ? 16 : s == 9 ? 16 : s == 10 ? 32 : s == 11 ? 32 : s == 12 ? 64
       : s == 13 ? 4 : s == 14 ? 16 : s == 15 ? 8 : s == 16 ? 32 : s == 17 ? 16 : 64;
}
__host__ __device__ constexpr int tx_h(int s) {
  return s == 0 ? 4 : s == 1 ? 8 : s == 2 ? 16 : s == 3 ? 32 : s == 4 ? 64 : s == 5 ? 8 : s == 6 ? 4
       : s == 7 ? 16 : s == 8 ? 8 : s == 9 ? 32 : s == 10 ? 16 : s == 11 ? 64 : s == 12 ? 32
       : s == 13 ? 16 : s == 14 ? 4 : s == 15 ? 32 : s == 16 ? 8 : s == 17 ? 64 : 16;
}
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// Forward cos bits per (log2 w - 2, log2 h - 2)
// (rust/crates/svtav1-dsp/src/fwd_txfm.rs:91 / :101).
__host__ __device__ constexpr int fwd_cb_col(int w, int h) {
  constexpr int8_t T[5][5] = {{13, 13, 13, 0, 0}, {13, 13, 13, 12, 0}, {13, 13, 13, 12, 13},
                              {0, 13, 13, 12, 13}, {0, 0, 13, 12, 13}};
  return T[ilog2c(w) - 2][ilog2c(h) - 2];
}
__host__ __device__ constexpr int fwd_cb_row(int w, int h) {
  constexpr int8_t T[5][5] = {{13, 13, 12, 0, 0}, {13, 13, 13, 12, 0}, {13, 13, 12, 13, 12},
                              {0, 12, 13, 12, 11}, {0, 0, 12, 11, 10}};
  return T[ilog2c(w) - 2][ilog2c(h) - 2];
}
// Forward per-pass shifts (fwd_txfm.rs:110): {left shift in, round-right
// after column pass, round-right after row pass}.
struct FwdShift { int s0, s1, s2; };
__host__ __device__ constexpr FwdShift fwd_shift(int w, int h) {
  return (w == 64 && h == 64) ? FwdShift{0, 2, 2}
       : (w == 32 && h == 64) ? FwdShift{0, 2, 2}
       : (w == 64 && h == 32) ? FwdShift{2, 4, 2}
       : (w == 16 && h == 64) ? FwdShift{0, 2, 0}
       : (w == 64 && h == 16) ? FwdShift{2, 4, 0}
       : (w == 4 && h == 4) ? FwdShift{2, 0, 0}
       : (w == 8 && h == 8) ? FwdShift{2, 1, 0}
       : (w == 16 && h == 16) ? FwdShift{2, 2, 0}
       : (w == 32 && h == 32) ? FwdShift{2, 4, 0}
       : ((w == 4 && h == 8) || (w == 8 && h == 4)) ? FwdShift{2, 1, 0}
       : ((w == 8 && h == 16) || (w == 16 && h == 8)) ? FwdShift{2, 2, 0}
       : ((w == 16 && h == 32) || (w == 32 && h == 16)) ? FwdShift{2, 4, 0}
       : ((w == 4 && h == 16) || (w == 16 && h == 4)) ? FwdShift{2, 1, 0}
       : FwdShift{2, 2, 0};  // 8x32 / 32x8
}
// Inverse shifts (inv_transforms shift table, inv_txfm.rs:1888): round-right
// after the row pass and after the column pass.
struct InvShift { int r0, r1; };
__host__ __device__ constexpr InvShift inv_shift(int w, int h) {
  return (w == h) ? InvShift{w == 4 ? 0 : (w == 8 ? 1 : 2), 4}
       : ((w == 2 * h) || (h == 2 * w)) ? InvShift{(w == 4 || h == 4) ? 0 : 1, 4}
       : InvShift{(w == 4 || h == 4) ? 1 : 2, 4};
}

// TxType -> (col kind, row kind, ud flip, lr flip); kinds 0 DCT, 1 ADST,
// 2 FLIPADST (ADST kernel, flip outside), 3 identity
// (rust/crates/svtav1-dsp/src/txfm_dispatch.rs:89,:99).
struct TypeCfg { int col, row, ud, lr; };
__host__ __device__ inline TypeCfg type_cfg(int t) {
  const int COL[16] = {0, 1, 0, 1, 2, 0, 2, 1, 2, 3, 0, 3, 1, 3, 2, 3};
  const int ROW[16] = {0, 0, 1, 1, 0, 2, 2, 2, 1, 3, 3, 0, 3, 1, 3, 2};
  TypeCfg c;
  c.col = COL[t];
  c.row = ROW[t];
  c.ud = (t == 4 || t == 6 || t == 8 || t == 14);
  c.lr = (t == 5 || t == 6 || t == 7 || t == 15);
  return c;
}
// A 1D kernel exists for (kind, n): ADST only up to 16 points.
__host__ __device__ constexpr bool kind_ok(int kind, int n) { return !((kind == 1 || kind == 2) && n > 16); }

__device__ __forceinline__ int32_t rshift_round(int32_t v, int bit) {
  return bit == 0 ? v : (int32_t)(((int64_t)v + (1 << (bit - 1))) >> bit);
}
// rshift_round for values known to satisfy |v| < 2^31 - 2^(bit - 1) (the
// fused kernels' bounded intermediates): the rounding add stays in 32 bits.
__device__ __forceinline__ int32_t rshift_round32(int32_t v, int bit) {
  return bit == 0 ? v : (v + (1 << (bit - 1))) >> bit;
}

// Raw buffer view of a u16 plane: loads take a 32-bit per-lane byte offset,
// a wave-uniform byte offset and an immediate, so they need no 64-bit
// address arithmetic.  Offsets must stay below 2^32 bytes (planes of
// < 2^31 samples) and are non-negative.
struct PlaneBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ explicit PlaneBuf(const uint16_t* p)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(p), (short)0, (int)0xffffffff, 0x00020000)) {}
  // sample at byte offset vbyte (per lane) + sbyte (uniform)
  __device__ __forceinline__ int32_t at(uint32_t vbyte, uint32_t sbyte) const {
    return (int32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)vbyte, (int)sbyte, 0);
  }
};

__device__ __forceinline__ int32_t rshift64r(int64_t v, int bit) {
  return (int32_t)((v + ((int64_t)1 << (bit - 1))) >> bit);
}

// ---------------------------------------------------------------------------
// 1D forward kernel dispatch on (N, cos_bit) compile-time, kind at run time.
// ---------------------------------------------------------------------------
template <int N, int CB>
__device__ __forceinline__ void fdct_n(int32_t* x) {
#define ZAV1_FDCT_PICK(NN)                                   \
  if constexpr (N == NN) {                                   \
    if constexpr (CB == 10) fdct##NN##_cb10(x);              \
    else if constexpr (CB == 11) fdct##NN##_cb11(x);         \
    else if constexpr (CB == 12) fdct##NN##_cb12(x);         \
    else fdct##NN##_cb13(x);                                 \
  }
  ZAV1_FDCT_PICK(4)
  ZAV1_FDCT_PICK(8)
  ZAV1_FDCT_PICK(16)
  ZAV1_FDCT_PICK(32)
  ZAV1_FDCT_PICK(64)
#undef ZAV1_FDCT_PICK
}

template <int N, int CB>
__device__ __forceinline__ void fdct_n_narrow(int32_t* x) {
#define ZAV1_FDCT_PICK(NN)                                   \
  if constexpr (N == NN) {                                   \
    if constexpr (CB == 10) fdct##NN##_cb10_n(x);            \
    else if constexpr (CB == 11) fdct##NN##_cb11_n(x);       \
    else if constexpr (CB == 12) fdct##NN##_cb12_n(x);       \
    else fdct##NN##_cb13_n(x);                               \
  }
  ZAV1_FDCT_PICK(4)
  ZAV1_FDCT_PICK(8)
  ZAV1_FDCT_PICK(16)
  ZAV1_FDCT_PICK(32)
  ZAV1_FDCT_PICK(64)
#undef ZAV1_FDCT_PICK
}

// Input-range bookkeeping for the narrow (24-bit operand, 32-bit sum)
// forward transforms.  The 4-point ADST's largest partial sum is
// y2 - y0 + y3 = (s4 - s1) x0 - (s1 + s2) x1 + s3 x2 + (s2 - s4) x3 with
// the products formed separately: at most (2 (s1 + s2 + s4) + s3) * B.
constexpr int kSinpiHost[4][5] = {{0, 330, 621, 836, 951}, {0, 660, 1241, 1672, 1901}, {0, 1321, 2482, 3344, 3803},
                                  {0, 2642, 4964, 6689, 7606}};
__host__ __device__ constexpr long long adst4_sum_gain(int cb) {
  return 2LL * (kSinpiHost[cb - 10][1] + kSinpiHost[cb - 10][2] + kSinpiHost[cb - 10][4]) + kSinpiHost[cb - 10][3];
}
__host__ __device__ constexpr long long identity_mult(int n) {
  return n == 4 ? 5793 : n == 8 ? 2 * 4096 : n == 16 ? 2 * 5793 : n == 32 ? 4 * 4096 : 4 * 5793;
}
// kind: 0 DCT, 1 ADST, 3 identity.  True when the narrow form of the 1D
// forward transform is exact for every input with |x| <= B.
__host__ __device__ constexpr bool fwd_narrow_ok(int kind, int n, int cb, long long B) {
  if (B < 0) return false;
  if (kind == 3) return B < (1LL << 23) && B * identity_mult(n) + 2048 < (1LL << 31);
  if (kind == 1 && n == 4) return 3 * B < (1LL << 23) && adst4_sum_gain(cb) * B + (1LL << (cb - 1)) < (1LL << 31);
  return B <= ztx_bound(kind == 0 ? 0 : 1, n, cb).bmax;
}
// Bound on the outputs of the 1D forward transform for inputs |x| <= B.
__host__ __device__ constexpr long long fwd_out_bound(int kind, int n, int cb, long long B) {
  if (kind == 3) return (B * identity_mult(n) + 2048) / 4096 + 1;
  if (kind == 1 && n == 4) return (adst4_sum_gain(cb) * B + (1LL << (cb - 1))) / (1LL << cb) + 1;
  const ZTxBound b = ztx_bound(kind == 0 ? 0 : 1, n, cb);
  return (long long)(b.g * (double)B + b.c) + 1;
}
// Input bound of the row pass of a w x h forward 2D transform whose column
// inputs are bounded by bcol (max over the column kinds the size allows).
__host__ __device__ constexpr long long fwd_row_bound(int h, int cbc, int s1, long long bcol) {
  long long m = fwd_out_bound(0, h, cbc, bcol);
  const long long mi = fwd_out_bound(3, h, cbc, bcol);
  if (mi > m) m = mi;
  if (h <= 16) {
    const long long ma = fwd_out_bound(1, h, cbc, bcol);
    if (ma > m) m = ma;
  }
  return s1 > 0 ? (m + (1LL << (s1 - 1))) >> s1 : m;
}
// Column-input bound of the fused kernels: residuals of planes of at most
// 10 bits, |r| <= 1023, shifted left by s0.
__host__ __device__ constexpr long long fwd_col_bound(int s0) { return 1023LL << s0; }

template <int N, int CB>
__device__ __forceinline__ void fadst4_narrow(int32_t* x) {
  const int32_t* s = kSinpi[CB - 10];
  const int32_t x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
  // (no early exit for an all-zero input: every output is then R >> CB = 0)
  const int32_t s0 = __mul24(s[1], x0), s1 = __mul24(s[4], x0), s2 = __mul24(s[2], x1);
  const int32_t s3 = __mul24(s[1], x1), s4 = __mul24(s[3], x2), s5 = __mul24(s[4], x3);
  const int32_t s6 = __mul24(s[2], x3);
  const int32_t s7 = x0 + x1 - x3;
  const int32_t y0 = s0 + s2 + s5, y1 = __mul24(s[3], s7), y2 = s1 - s3 + s6, y3 = s4;
  constexpr int32_t R = 1 << (CB - 1);
  x[0] = (y0 + y3 + R) >> CB;
  x[1] = (y1 + R) >> CB;
  x[2] = (y2 - y3 + R) >> CB;
  x[3] = (y2 - y0 + y3 + R) >> CB;
}

template <int N, int CB>
__device__ __forceinline__ void fadst_n_narrow(int32_t* x) {
  if constexpr (N == 4) {
    fadst4_narrow<N, CB>(x);
  } else if constexpr (N == 8) {
    if constexpr (CB == 10) fadst8_cb10_n(x);
    else if constexpr (CB == 11) fadst8_cb11_n(x);
    else if constexpr (CB == 12) fadst8_cb12_n(x);
    else fadst8_cb13_n(x);
  } else if constexpr (N == 16) {
    if constexpr (CB == 10) fadst16_cb10_n(x);
    else if constexpr (CB == 11) fadst16_cb11_n(x);
    else if constexpr (CB == 12) fadst16_cb12_n(x);
    else fadst16_cb13_n(x);
  }
}

template <int N>
__device__ __forceinline__ void fidentity_narrow(int32_t* x) {
#pragma unroll
  for (int i = 0; i < N; i++) {
    if constexpr (N == 8) x[i] = x[i] * 2;
    else if constexpr (N == 32) x[i] = x[i] * 4;
    else x[i] = (__mul24(x[i], (int32_t)identity_mult(N)) + 2048) >> 12;
  }
}

template <int N, int CB>
__device__ __forceinline__ void fadst_n(int32_t* x) {
  if constexpr (N == 4) {
    const int32_t* s = kSinpi[CB - 10];
    const int32_t x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
    if ((x0 | x1 | x2 | x3) == 0) return;
    const int64_t s0 = (int64_t)s[1] * x0, s1 = (int64_t)s[4] * x0, s2 = (int64_t)s[2] * x1;
    const int64_t s3 = (int64_t)s[1] * x1, s4 = (int64_t)s[3] * x2, s5 = (int64_t)s[4] * x3;
    const int64_t s6 = (int64_t)s[2] * x3;
    const int64_t s7 = (int64_t)(int32_t)(x0 + x1) - x3;
    const int64_t y0 = s0 + s2 + s5, y1 = (int64_t)s[3] * s7, y2 = s1 - s3 + s6, y3 = s4;
    x[0] = rshift64r(y0 + y3, CB);
    x[1] = rshift64r(y1, CB);
    x[2] = rshift64r(y2 - y3, CB);
    x[3] = rshift64r(y2 - y0 + y3, CB);
  } else if constexpr (N == 8) {
    if constexpr (CB == 10) fadst8_cb10(x);
    else if constexpr (CB == 11) fadst8_cb11(x);
    else if constexpr (CB == 12) fadst8_cb12(x);
    else fadst8_cb13(x);
  } else if constexpr (N == 16) {
    if constexpr (CB == 10) fadst16_cb10(x);
    else if constexpr (CB == 11) fadst16_cb11(x);
    else if constexpr (CB == 12) fadst16_cb12(x);
    else fadst16_cb13(x);
  }
}

// identity scalings: x sqrt2 (4), x2 (8), x2sqrt2 (16), x4 (32), x4sqrt2 (64)
template <int N>
__device__ __forceinline__ void fidentity_n(int32_t* x) {
#pragma unroll
  for (int i = 0; i < N; i++) {
    if constexpr (N == 4) x[i] = rshift64r((int64_t)x[i] * 5793, 12);
    else if constexpr (N == 8) x[i] = x[i] * 2;
    else if constexpr (N == 16) x[i] = rshift64r((int64_t)x[i] * 2 * 5793, 12);
    else if constexpr (N == 32) x[i] = x[i] * 4;
    else x[i] = rshift64r((int64_t)x[i] * 4 * 5793, 12);
  }
}

// B >= 0: every input satisfies |x| <= B, and each kind whose narrow form is
// exact at that bound uses it; B < 0 (no bound known): 64-bit butterflies.
template <int N, int CB, long long B = -1>
__device__ __forceinline__ void fwd1d(int32_t* x, int kind) {
  if (kind == 0) {
    if constexpr (fwd_narrow_ok(0, N, CB, B)) fdct_n_narrow<N, CB>(x);
    else fdct_n<N, CB>(x);
  } else if (kind == 3) {
    if constexpr (fwd_narrow_ok(3, N, CB, B)) fidentity_narrow<N>(x);
    else fidentity_n<N>(x);
  } else {
    if constexpr (N <= 16) {
      if constexpr (fwd_narrow_ok(1, N, CB, B)) fadst_n_narrow<N, CB>(x);
      else fadst_n<N, CB>(x);
    }
  }
}

// 1D inverse, INV_COS_BIT 12, stage clamp `rng`.
template <int N>
__device__ __forceinline__ void inv1d(int32_t* x, int kind, int rng) {
  if (kind == 0) {
    if constexpr (N == 4) idct4(x, rng);
    else if constexpr (N == 8) idct8(x, rng);
    else if constexpr (N == 16) idct16(x, rng);
    else if constexpr (N == 32) idct32(x, rng);
    else idct64(x, rng);
  } else if (kind == 3) {
    fidentity_n<N>(x);  // iidentity scalings equal the forward ones
  } else {
    if constexpr (N == 4) {
      const int32_t* s = kSinpi[2];
      const int32_t x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3];
      if ((x0 | x1 | x2 | x3) == 0) return;
      // i32 products with two's-complement wrap, as the reference computes them
      auto m = [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); };
      int32_t s0 = m(s[1], x0), s1 = m(s[2], x0), s2 = m(s[3], x1), s3 = m(s[4], x2);
      int32_t s4 = m(s[1], x2), s5 = m(s[2], x3), s6 = m(s[4], x3);
      int32_t s7 = (int32_t)((uint32_t)(x0 - x2) + (uint32_t)x3);
      s0 = (int32_t)((uint32_t)s0 + (uint32_t)s3);
      s1 = (int32_t)((uint32_t)s1 - (uint32_t)s4);
      s3 = s2;
      s2 = m(s[3], s7);
      s0 = (int32_t)((uint32_t)s0 + (uint32_t)s5);
      s1 = (int32_t)((uint32_t)s1 - (uint32_t)s6);
      const int32_t y0 = (int32_t)((uint32_t)s0 + (uint32_t)s3);
      const int32_t y1 = (int32_t)((uint32_t)s1 + (uint32_t)s3);
      const int32_t y2 = s2;
      const int32_t y3 = (int32_t)((uint32_t)s0 + (uint32_t)s1 - (uint32_t)s3);
      x[0] = rshift64r(y0, 12);
      x[1] = rshift64r(y1, 12);
      x[2] = rshift64r(y2, 12);
      x[3] = rshift64r(y3, 12);
    } else if constexpr (N == 8) {
      iadst8(x, rng);
    } else if constexpr (N == 16) {
      iadst16(x, rng);
    }
  }
}

// log_scale per TxSize (rust/crates/svtav1-encoder/src/quant.rs:238).
__host__ __device__ constexpr int tx_log_scale(int w, int h) {
  return (w * h > 1024) ? 2 : ((w * h > 256) ? 1 : 0);
}

// Sum / max over aligned groups of G = 1, 2, 4, 8 or 16 lanes, every lane
// of a group receiving the result, through DPP lane permutations (no LDS
// crossbar): G = 16 rotates within the 16-lane row by 8, 4, 2, 1; G = 8
// mirrors the half-row, then swaps lane pairs within quads; G = 4 swaps
// within quads; G = 2 swaps neighbours.
template <int G>
__device__ __forceinline__ int32_t dpp_step(int32_t v, int k) {
  // k-th permutation of the group's reduction
  if constexpr (G == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
  } else if constexpr (G == 16) {
    return k == 0 ? __builtin_amdgcn_update_dpp(0, v, 0x128, 0xf, 0xf, false)
         : k == 1 ? __builtin_amdgcn_update_dpp(0, v, 0x124, 0xf, 0xf, false)
         : k == 2 ? __builtin_amdgcn_update_dpp(0, v, 0x122, 0xf, 0xf, false)
                  : __builtin_amdgcn_update_dpp(0, v, 0x121, 0xf, 0xf, false);
  } else if constexpr (G == 8) {
    return k == 0 ? __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false)
         : k == 1 ? __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false)
                  : __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
  } else {
    return k == 0 ? __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false)
                  : __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
  }
}
template <int G>
constexpr int dpp_steps() {
  return G == 16 ? 4 : (G == 8 ? 3 : (G == 4 ? 2 : (G == 2 ? 1 : 0)));
}
template <int G>
__device__ __forceinline__ int32_t dpp_sum_i32(int32_t v) {
  static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group");
  constexpr int NS = dpp_steps<G>();
#pragma unroll
  for (int k = 0; k < NS; k++) v += dpp_step<G>(v, k);
  return v;
}
template <int G>
__device__ __forceinline__ int32_t dpp_max_i32(int32_t v) {
  constexpr int NS = dpp_steps<G>();
#pragma unroll
  for (int k = 0; k < NS; k++) v = max(v, dpp_step<G>(v, k));
  return v;
}
// For per-lane values below 2^52: the low 24 bits and the rest are summed
// as two 32-bit values (16 lanes of each stay below 2^32), so every step is
// one fused 32-bit DPP add.
template <int G>
__device__ __forceinline__ uint64_t dpp_sum_u64(uint64_t v) {
  const int32_t lo = dpp_sum_i32<G>((int32_t)((uint32_t)v & 0xffffffu));
  const int32_t hi = dpp_sum_i32<G>((int32_t)(uint32_t)(v >> 24));
  return ((uint64_t)(uint32_t)hi << 24) + (uint32_t)lo;
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

// The first k passes of the reference's exchange sort
// (c_exchange_sort_by, rust/crates/svtav1-encoder/src/leaf_funnel/
// rate_tables.rs:108: for position i, every later position j whose cost is
// strictly below the running minimum swaps with position i), lane-parallel
// in one wave: lane j holds position j's candidate `idx` and its `cost`
// (n <= 64 lanes live, the rest hold ~0 costs and never move).  Pass i's
// swaps move the running-minimum records: with the records at positions
// i = r0 < r1 < ... < rm (the positions whose cost is below every cost in
// [i, position)), position i ends with rm's candidate and each r(q) with
// r(q-1)'s -- the sequential scan's result, found by an exclusive prefix
// minimum over the lanes and a ballot.  After pass i, position i is final;
// `out[i]` (lane-uniform) receives it.  Wave-uniform control flow.
__device__ inline void wave_exchange_first_k(int& idx, uint64_t& cost, int n, int k, int lane, uint8_t* out,
                                             uint64_t* out_cost) {
  for (int i = 0; i < k; i++) {
    const bool live = lane >= i && lane < n;
    const uint64_t c = live ? cost : ~0ull;
    // exclusive prefix minimum over lanes i .. lane-1
    uint64_t pre = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t lo = __shfl_up((uint32_t)pre, d, 64), hi = __shfl_up((uint32_t)(pre >> 32), d, 64);
      const uint64_t o = ((uint64_t)hi << 32) | lo;
      if (lane >= d) pre = o < pre ? o : pre;
    }
    uint64_t ex;
    {
      const uint32_t lo = __shfl_up((uint32_t)pre, 1, 64), hi = __shfl_up((uint32_t)(pre >> 32), 1, 64);
      ex = lane >= 1 ? (((uint64_t)hi << 32) | lo) : ~0ull;
    }
    const bool rec = live && (lane == i || c < ex);
    const uint64_t mask = __ballot(rec);
    const int last = 63 - __builtin_clzll(mask);
    // the previous record below this lane
    const uint64_t below = mask & ((lane ? (~0ull >> (64 - lane)) : 0ull));
    const int prev = below ? 63 - __builtin_clzll(below) : lane;
    const int from = lane == i ? last : (rec ? prev : lane);
    const int ni = __shfl(idx, from, 64);
    const uint32_t clo = __shfl((uint32_t)cost, from, 64), chi = __shfl((uint32_t)(cost >> 32), from, 64);
    idx = ni;
    cost = ((uint64_t)chi << 32) | clo;
    const int fi = __shfl(idx, i, 64);
    const uint32_t flo = __shfl((uint32_t)cost, i, 64), fhi = __shfl((uint32_t)(cost >> 32), i, 64);
    if (lane == 0) {
      out[i] = (uint8_t)fi;
      if (out_cost) out_cost[i] = ((uint64_t)fhi << 32) | flo;
    }
  }
}

}  // namespace zav1
