// mlp.hpp — the feature-prediction network of the learned narrow phase.
//
// 18 canonical-frame coordinates (padded to 32) -> 64 -> 64 -> 32 outputs,
// GELU (hard-sigmoid form) between layers, bf16 weights and activations, FP32
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

// GELU in the hard-sigmoid form x * clamp(1.702 x / 6 + 1/2, 0, 1): the
// sigmoid of x * sigmoid(1.702 x) (Hendrycks & Gimpel 2016) replaced by its
// piecewise-linear hard-sigmoid, one fma, one med3 and one mul per unit.
constexpr float kHardA = 1.702f / 6.0f;
__device__ inline float gelu(float x) {
  return x * __builtin_amdgcn_fmed3f(kHardA * x + 0.5f, 0.0f, 1.0f);
}
__device__ inline float gelu_grad(float x) {
  const float t = kHardA * x + 0.5f;
  return (t <= 0.0f) ? 0.0f : ((t >= 1.0f) ? 1.0f : 2.0f * kHardA * x + 0.5f);
}

__device__ inline v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// four FP32 -> bf16 (round to nearest even, v_cvt_pk_bf16_f32)
__device__ inline v4s pack4(float a, float b, float c, float d) {
  const v4bf t = {static_cast<__bf16>(a), static_cast<__bf16>(b), static_cast<__bf16>(c),
                  static_cast<__bf16>(d)};
  return __builtin_bit_cast(v4s, t);
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
