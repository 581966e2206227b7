// Hadamard building blocks of the aom SATD kernels for gfx950 (device
// code shared by the SATD batch and the fused MDS0 kernel): 8x8 / 4x4 tile
// transforms with the reference's 16-bit store rules, and the residual
// readers they take.
//
//   aom_hadamard_4x4 / 8x8 (+ 16x16 / 32x32 combine rules)
//   rust/crates/svtav1-dsp/src/hadamard.rs:643-948
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zav1 {

__device__ __forceinline__ void had8_col(const int32_t* s, int16_t* o) {
  // 8-point butterfly on s[0..7]; C output permutation, 16-bit store.
  const int32_t b0 = s[0] + s[1], b1 = s[0] - s[1], b2 = s[2] + s[3], b3 = s[2] - s[3];
  const int32_t b4 = s[4] + s[5], b5 = s[4] - s[5], b6 = s[6] + s[7], b7 = s[6] - s[7];
  const int32_t c0 = b0 + b2, c1 = b1 + b3, c2 = b0 - b2, c3 = b1 - b3;
  const int32_t c4 = b4 + b6, c5 = b5 + b7, c6 = b4 - b6, c7 = b5 - b7;
  o[0] = (int16_t)(c0 + c4); o[7] = (int16_t)(c1 + c5); o[3] = (int16_t)(c2 + c6);
  o[4] = (int16_t)(c3 + c7); o[2] = (int16_t)(c0 - c4); o[6] = (int16_t)(c1 - c5);
  o[1] = (int16_t)(c2 - c6); o[5] = (int16_t)(c3 - c7);
}

// Residual readers: a(r, c) is the i16 residual at row r, column c of a tile.
// Packed: a stored i16 residual.  Diff: source minus prediction in 16 bits
// (the reference's `src as i16 - pred as i16`).
struct PackedRes {
  const int16_t* p;
  int st;
  __device__ __forceinline__ int16_t operator()(int r, int c) const { return p[r * st + c]; }
  __device__ __forceinline__ PackedRes at(int r0, int c0) const { return {p + r0 * st + c0, st}; }
};
template <typename Pix>
struct DiffRes {
  const Pix* s;
  const uint16_t* q;
  int ss, qs;
  __device__ __forceinline__ int16_t operator()(int r, int c) const {
    return (int16_t)((int16_t)s[r * ss + c] - (int16_t)q[r * qs + c]);
  }
  __device__ __forceinline__ DiffRes at(int r0, int c0) const { return {s + r0 * ss + c0, q + r0 * qs + c0, ss, qs}; }
};

// 8x8 tile of residual reader `a` -> 64 coefficients (i16-valued).
template <class R>
__device__ __forceinline__ void had8_tile(const R& a, int32_t* out) {
  int16_t buf[64];
  int32_t col[8];
#pragma unroll
  for (int c = 0; c < 8; c++) {
#pragma unroll
    for (int r = 0; r < 8; r++) col[r] = a(r, c);
    had8_col(col, buf + c * 8);
  }
  int16_t o[8];
#pragma unroll
  for (int idx = 0; idx < 8; idx++) {
#pragma unroll
    for (int k = 0; k < 8; k++) col[k] = buf[idx + 8 * k];
    had8_col(col, o);
#pragma unroll
    for (int k = 0; k < 8; k++) out[idx * 8 + k] = o[k];
  }
}

template <class R>
__device__ __forceinline__ void had4_tile(const R& a, int32_t* out) {
  int16_t b1[16], b2[16];
  auto col4 = [](const int16_t* s, int stride, int16_t* o) {
    const int16_t x0 = (int16_t)((int16_t)(s[0] + s[stride]) >> 1);
    const int16_t x1 = (int16_t)((int16_t)(s[0] - s[stride]) >> 1);
    const int16_t x2 = (int16_t)((int16_t)(s[2 * stride] + s[3 * stride]) >> 1);
    const int16_t x3 = (int16_t)((int16_t)(s[2 * stride] - s[3 * stride]) >> 1);
    o[0] = (int16_t)(x0 + x2); o[1] = (int16_t)(x1 + x3);
    o[2] = (int16_t)(x0 - x2); o[3] = (int16_t)(x1 - x3);
  };
  int16_t tile[16];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) tile[r * 4 + c] = a(r, c);
#pragma unroll
  for (int i = 0; i < 4; i++) col4(tile + i, 4, b1 + 4 * i);
#pragma unroll
  for (int i = 0; i < 4; i++) col4(b1 + i, 4, b2 + 4 * i);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) out[i * 4 + j] = b2[j * 4 + i];
}

__device__ __forceinline__ int16_t sat16(int32_t v) {
  return (int16_t)(v < -32768 ? -32768 : (v > 32767 ? 32767 : v));
}

}  // namespace zav1
