// mlp.hpp — the feature-prediction network of the learned narrow phase.
//
// 18 canonical-frame coordinates (padded to 32) -> 64 -> 64 -> 32 outputs,
// ReLU between layers, bf16 weights and activations, FP32
// accumulation on MFMA (v_mfma_f32_16x16x16_bf16).  Outputs: [0,15) logits
// of the closest feature pair (0..8 edge k of A x edge m of B, 9..14 vertex-
// face), 15 the exact-path logit, [16,20) barycentric parameters of P and Q.
//
// The layers are evaluated transposed, Y^T = W^T X^T: with pairs on the MFMA
// N dimension, the accumulator layout of one layer (lane l holds rows
// 4*(l/16)+j, column l%16) is exactly the B-operand layout the next layer
// needs, so activations never leave registers between layers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dh {
namespace mlp {

constexpr int kIn = 32;    // 18 used
constexpr int kInUsed = 18;
constexpr int kHid = 64;
constexpr int kOut = 32;   // 20 used
constexpr int kNumFeat = 15;
constexpr int kFlagOut = 15;
constexpr int kBaryOut = 16;

// flat FP32 parameter layout
constexpr int kW1 = 0;                     // [kIn][kHid]
constexpr int kB1 = kW1 + kIn * kHid;      // [kHid]
constexpr int kW2 = kB1 + kHid;            // [kHid][kHid]
constexpr int kB2 = kW2 + kHid * kHid;     // [kHid]
constexpr int kW3 = kB2 + kHid;            // [kHid][kOut]
constexpr int kB3 = kW3 + kHid * kOut;     // [kOut]
constexpr int kNumParams = kB3 + kOut;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> bf16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t h) {
  const uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// ReLU between layers.  On bf16 values it is a 16-bit integer max against 0
// (a negative bf16, -0 included, is a negative int16), so relu4 applies it to
// four packed activations with two v_pk_max_i16.
__device__ inline float relu(float x) { return x > 0.0f ? x : 0.0f; }
__device__ inline float relu_grad(float x) { return x > 0.0f ? 1.0f : 0.0f; }

__device__ inline v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// v_mfma_f32_16x16x32_bf16 (gfx950): lane l holds 8 A values of row l%16 and
// 8 B values of column l%16; A slot (l/16, j) multiplies B slot (l/16, j), so
// any assignment of the 32 k indices to (l/16, j) slots works as long as A
// and B use the same one.
__device__ inline v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                 __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

__device__ inline v8s cat8(v4s lo, v4s hi) {
  return {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// four FP32 -> bf16 (round to nearest even, v_cvt_pk_bf16_f32)
__device__ inline v4s pack4(float a, float b, float c, float d) {
  const v4bf t = {static_cast<__bf16>(a), static_cast<__bf16>(b), static_cast<__bf16>(c),
                  static_cast<__bf16>(d)};
  return __builtin_bit_cast(v4s, t);
}

__device__ inline v4s relu4(v4s x) {
  const v4s zero = {0, 0, 0, 0};
  return __builtin_elementwise_max(x, zero);
}

__device__ inline uint16_t bf16_bits(float f) {
  const __bf16 h = static_cast<__bf16>(f);
  return __builtin_bit_cast(uint16_t, h);
}

// A-operand tile of W^T for out-tile `to`, k-tile `tk` of a [K][N] weight:
// lane l holds W[tk*16 + 4*(l/16) + j][to*16 + l%16], j = 0..3.
__device__ inline v4s load_wt(const float* W, int N, int to, int tk, int lane) {
  const int n = to * 16 + (lane & 15);
  const int k = tk * 16 + 4 * (lane >> 4);
  return pack4(W[(k + 0) * N + n], W[(k + 1) * N + n], W[(k + 2) * N + n], W[(k + 3) * N + n]);
}

// A-operand tile of W (not transposed) for backpropagation through a [K][N]
// weight: rows are K (units of the layer input), columns N (its outputs):
// lane l holds W[tk*16 + l%16][tn*16 + 4*(l/16) + j].
__device__ inline v4s load_w(const float* W, int N, int tk, int tn, int lane) {
  const int k = tk * 16 + (lane & 15);
  const int n = tn * 16 + 4 * (lane >> 4);
  return pack4(W[k * N + n], W[k * N + n + 1], W[k * N + n + 2], W[k * N + n + 3]);
}

}  // namespace mlp
}  // namespace dh
