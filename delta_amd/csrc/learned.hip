// learned.hip — the learned narrow phase: labels from the FP64 brute force,
// MFMA training of the feature network, and the fused inference kernel.
//
// Inference per pair: canonical-frame inputs -> network (bf16 MFMA) -> the
// most likely edge-edge and the most likely vertex-face feature -> both
// evaluated in binary64 with bfSolver's arithmetic and compared by bfSolver's
// rule -> exact_choice() (pair_features.hpp) proves that the winner is the
// feature bfSolver itself selects, or names the few tied features of a
// degenerate closest pair, which are then evaluated and folded in bfSolver's
// order.  The output of every accepted pair is bfSolver's (P, Q, d) bit for
// bit.  The rest (intersecting pairs, wrong predictions, pairs the network
// marks for the exact path) are compacted into per-workgroup queues that the
// FP64 brute force drains.  The network therefore decides only how much work
// a pair costs, never its result.
#include "exact_math.hpp"
#include "mlp.hpp"
#include "pair_features.hpp"
#include "runtime.hpp"

#include <vector>

namespace dh {

using namespace mlp;

__device__ __forceinline__ Tri<double> load_tri_soa(const double* __restrict__ soa, int64_t n,
                                                    int64_t i, int base) {
  Tri<double> t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.v[k].x = soa[(base + k) * n + i];
    t.v[k].y = soa[(base + 3 + k) * n + i];
    t.v[k].z = soa[(base + 6 + k) * n + i];
  }
  return t;
}

// ------------------------------------------------------------------ labels
__global__ __launch_bounds__(256) void k_labels(const double* __restrict__ soa, int64_t n,
                                                uint16_t* __restrict__ X,
                                                uint8_t* __restrict__ feat,
                                                float* __restrict__ bary,
                                                float* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Tri<double> A = load_tri_soa(soa, n, i, 0), B = load_tri_soa(soa, n, i, 9);
  float in[18];
  const bool frame = canonical_inputs(A, B, in);
  uint16_t* x = X + i * kIn;
  for (int k = 0; k < kIn; ++k) x[k] = (k < kInUsed) ? bf16_bits(in[k]) : 0;
  V3<double> P, Q;
  int ssid = 0, vid = 0;
  const bool edge = bf_solver_ids(A, B, P, Q, ssid, vid);
  const int code = edge ? ssid : 9 + vid;
  feat[i] = static_cast<uint8_t>(code);
  double al, be, ga, de;
  barycentric(A, P, al, be);
  barycentric(B, Q, ga, de);
  auto clampf = [](double v) { return static_cast<float>(v < -0.5 ? -0.5 : (v > 1.5 ? 1.5 : v)); };
  bary[4 * i + 0] = clampf(al);
  bary[4 * i + 1] = clampf(be);
  bary[4 * i + 2] = clampf(ga);
  bary[4 * i + 3] = clampf(de);
  // exact path: bfSolver's own winner does not pass the support check (an
  // intersecting or touching pair, a near-parallel or near-coplanar closest
  // pair, a degenerate triangle)
  V3<double> Pw, Qw;
  unsigned sA, sB;
  if (edge)
    eval_ee_value(A, B, code, Pw, Qw, sA, sB);
  else
    eval_vf_value(A, B, code, Pw, Qw, sA, sB);
  flag[i] = (frame && support_choice(A, B, P, Q, sA, sB, code) != kUnresolvedAll) ? 0.0f : 1.0f;
}

// ------------------------------------------------------------- forward core
struct FwdWeights {
  v4s w1[4][2];  // W1^T tiles [out-tile][k-tile]
  v4s w2[4][4];
  v4s w3[2][4];
};

__device__ inline void load_fwd(const float* __restrict__ W, FwdWeights& fw, int lane) {
#pragma unroll
  for (int to = 0; to < 4; ++to)
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) fw.w1[to][tk] = load_wt(W + kW1, kHid, to, tk, lane);
#pragma unroll
  for (int to = 0; to < 4; ++to)
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) fw.w2[to][tk] = load_wt(W + kW2, kHid, to, tk, lane);
#pragma unroll
  for (int to = 0; to < 2; ++to)
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) fw.w3[to][tk] = load_wt(W + kW3, kOut, to, tk, lane);
}

// x_b[tk]: B operand of layer 1 (inputs k = tk*16 + 4*(lane/16) + j of pair
// lane%16).  Produces pre-activations z1, z2 and outputs z3 (D layout).
__device__ inline void forward16(const FwdWeights& fw, const float* __restrict__ W,
                                 const v4s x_b[2], int lane, v4f z1[4], v4f z2[4], v4f z3[2]) {
  const int g = lane >> 4;
  v4s a1[4], a2[4];
#pragma unroll
  for (int to = 0; to < 4; ++to) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(fw.w1[to][0], x_b[0], acc);
    acc = mfma16(fw.w1[to][1], x_b[1], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] += W[kB1 + to * 16 + 4 * g + j];
    z1[to] = acc;
    a1[to] = relu4(pack4(acc[0], acc[1], acc[2], acc[3]));
  }
#pragma unroll
  for (int to = 0; to < 4; ++to) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) acc = mfma16(fw.w2[to][tk], a1[tk], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] += W[kB2 + to * 16 + 4 * g + j];
    z2[to] = acc;
    a2[to] = relu4(pack4(acc[0], acc[1], acc[2], acc[3]));
  }
#pragma unroll
  for (int to = 0; to < 2; ++to) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) acc = mfma16(fw.w3[to][tk], a2[tk], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] += W[kB3 + to * 16 + 4 * g + j];
    z3[to] = acc;
  }
}

__device__ inline v4s x_operand(const uint16_t* __restrict__ xrow, int tk, int lane) {
  const uint16_t* p = xrow + tk * 16 + 4 * (lane >> 4);
  v4s r;
  r[0] = static_cast<short>(p[0]);
  r[1] = static_cast<short>(p[1]);
  r[2] = static_cast<short>(p[2]);
  r[3] = static_cast<short>(p[3]);
  return r;
}

// Forward pass of n pairs (outputs [n][kOut] FP32); one wave per 16 pairs.
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ W,
                                                    const uint16_t* __restrict__ X, int64_t n,
                                                    float* __restrict__ out) {
  const int lane = threadIdx.x;
  FwdWeights fw;
  load_fwd(W, fw, lane);
  const int64_t ntiles = (n + 15) / 16;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t pair = t * 16 + (lane & 15);
    const bool valid = pair < n;
    v4s xb[2];
    const uint16_t zero[kIn] = {};
    const uint16_t* xrow = valid ? X + pair * kIn : zero;
    xb[0] = x_operand(xrow, 0, lane);
    xb[1] = x_operand(xrow, 1, lane);
    v4f z1[4], z2[4], z3[2];
    forward16(fw, W, xb, lane, z1, z2, z3);
    if (valid)
      for (int to = 0; to < 2; ++to)
        for (int j = 0; j < 4; ++j) out[pair * kOut + to * 16 + 4 * (lane >> 4) + j] = z3[to][j];
  }
}

// -------------------------------------------------------------- training
// Loss per pair: cross-entropy over the 15 feature logits + 0.5 * binary
// cross-entropy of the exact-path logit + mean squared error of the four
// barycentric outputs; averaged over the batch.
struct TrainArgs {
  const float* W;
  const uint16_t* X;
  const uint8_t* feat;
  const float* bary;
  const float* flag;
  int64_t offset;   // first pair of the batch
  int batch;        // pairs in the batch
  float* grad;      // [kNumParams], accumulated
  float* loss;      // [1], accumulated
};

constexpr int kTrainPairs = 64;  // pairs per wave iteration (4 MFMA tiles)

__global__ __launch_bounds__(64) void k_mlp_train(TrainArgs a) {
  __shared__ uint16_t x_s[kTrainPairs][kIn];
  __shared__ uint16_t a1_s[kHid][kTrainPairs];
  __shared__ uint16_t a2_s[kHid][kTrainPairs];
  __shared__ uint16_t d1_s[kHid][kTrainPairs];
  __shared__ uint16_t d2_s[kHid][kTrainPairs];
  __shared__ uint16_t d3_s[kOut][kTrainPairs];
  __shared__ float g_s[kNumParams];
  const int lane = threadIdx.x;
  const int g = lane >> 4, c = lane & 15;
  const float* W = a.W;
  FwdWeights fw;
  load_fwd(W, fw, lane);
  v4s w3b[4][2], w2b[4][4];
#pragma unroll
  for (int tk = 0; tk < 4; ++tk)
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) w3b[tk][tn] = load_w(W + kW3, kOut, tk, tn, lane);
#pragma unroll
  for (int tk = 0; tk < 4; ++tk)
#pragma unroll
    for (int tn = 0; tn < 4; ++tn) w2b[tk][tn] = load_w(W + kW2, kHid, tk, tn, lane);
  for (int i = lane; i < kNumParams; i += 64) g_s[i] = 0.f;
  float loss_acc = 0.f;
  const float inv_b = 1.0f / static_cast<float>(a.batch);
  const int nchunks = (a.batch + kTrainPairs - 1) / kTrainPairs;
  __syncthreads();
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int base = ch * kTrainPairs;
    {  // stage the chunk's inputs (one pair per lane)
      const int p = base + lane;
      const bool v = p < a.batch;
      const uint16_t* src = a.X + (a.offset + p) * kIn;
      for (int k = 0; k < kIn; ++k) x_s[lane][k] = v ? src[k] : 0;
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
      const int pl = s * 16 + c;  // pair within the chunk
      const int p = base + pl;
      const bool valid = p < a.batch;
      v4s xb[2] = {x_operand(&x_s[pl][0], 0, lane), x_operand(&x_s[pl][0], 1, lane)};
      v4f z1[4], z2[4], z3[2];
      forward16(fw, W, xb, lane, z1, z2, z3);
      // activations for the weight gradients, [unit][pair]
#pragma unroll
      for (int to = 0; to < 4; ++to)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          a1_s[to * 16 + 4 * g + j][pl] = bf16_bits(relu(z1[to][j]));
          a2_s[to * 16 + 4 * g + j][pl] = bf16_bits(relu(z2[to][j]));
        }
      // loss and dL/dz3 (tile 0: logits 0..14 and the flag 15; tile 1: 16..19)
      const int64_t gp = a.offset + (valid ? p : 0);
      const int label = a.feat[gp];
      float mx = -3.0e38f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < kNumFeat) mx = fmaxf(mx, z3[0][j]);
      mx = fmaxf(mx, __shfl_xor(mx, 16));
      mx = fmaxf(mx, __shfl_xor(mx, 32));
      float se = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < kNumFeat) se += expf(z3[0][j] - mx);
      se += __shfl_xor(se, 16);
      se += __shfl_xor(se, 32);
      v4f dz3[2];
      float lp = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * g + j;
        float d = 0.f;
        if (o < kNumFeat) {
          const float sm = expf(z3[0][j] - mx) / se;
          d = sm - (o == label ? 1.f : 0.f);
          if (o == label) lp += -(z3[0][j] - mx - logf(se));
        } else {  // exact-path logit
          const float y = a.flag[gp];
          const float zz = z3[0][j];
          const float sg = 1.f / (1.f + expf(-zz));
          d = 0.5f * (sg - y);
          lp += 0.5f * (fmaxf(zz, 0.f) - zz * y + logf(1.f + expf(-fabsf(zz))));
        }
        dz3[0][j] = valid ? d * inv_b : 0.f;
        const int o1 = 16 + 4 * g + j;
        float d1 = 0.f;
        if (o1 < kBaryOut + 4) {
          const float r = z3[1][j] - a.bary[4 * gp + (o1 - kBaryOut)];
          d1 = 0.5f * r;  // d/dz of mean over 4 of r^2
          lp += 0.25f * r * r;
        }
        dz3[1][j] = valid ? d1 * inv_b : 0.f;
      }
      if (valid) loss_acc += lp;
      // backward through the layers (activations stay in registers)
      v4s dz3b[2];
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) {
        dz3b[tn] = pack4(dz3[tn][0], dz3[tn][1], dz3[tn][2], dz3[tn][3]);
#pragma unroll
        for (int j = 0; j < 4; ++j) d3_s[tn * 16 + 4 * g + j][pl] = bf16_bits(dz3[tn][j]);
      }
      v4s dz2b[4];
#pragma unroll
      for (int tu = 0; tu < 4; ++tu) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(w3b[tu][0], dz3b[0], acc);
        acc = mfma16(w3b[tu][1], dz3b[1], acc);
        float d[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          d[j] = acc[j] * relu_grad(z2[tu][j]);
          d2_s[tu * 16 + 4 * g + j][pl] = bf16_bits(d[j]);
        }
        dz2b[tu] = pack4(d[0], d[1], d[2], d[3]);
      }
#pragma unroll
      for (int tu = 0; tu < 4; ++tu) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tn = 0; tn < 4; ++tn) acc = mfma16(w2b[tu][tn], dz2b[tn], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          d1_s[tu * 16 + 4 * g + j][pl] = bf16_bits(acc[j] * relu_grad(z1[tu][j]));
      }
    }
    __syncthreads();
    // weight gradients of the chunk: dW[k][n] = sum_p act[k][p] dz[n][p]
    auto wgrad = [&](const uint16_t* act, int act_stride, bool act_is_x, const uint16_t* dz,
                     int K, int N, int woff, int boff) {
      for (int tk = 0; tk < K / 16; ++tk)
        for (int tn = 0; tn < N / 16; ++tn) {
          v4f acc = {0.f, 0.f, 0.f, 0.f};
          for (int ps = 0; ps < 4; ++ps) {
            const int p0 = ps * 16 + 4 * g;
            v4s A, Bm;
            for (int j = 0; j < 4; ++j) {
              const int k = tk * 16 + c;
              A[j] = static_cast<short>(act_is_x ? x_s[p0 + j][k] : act[k * act_stride + p0 + j]);
              Bm[j] = static_cast<short>(dz[(tn * 16 + c) * kTrainPairs + p0 + j]);
            }
            acc = mfma16(A, Bm, acc);
          }
          for (int j = 0; j < 4; ++j)
            atomicAdd(&g_s[woff + (tk * 16 + 4 * g + j) * N + tn * 16 + c], acc[j]);
        }
      // bias gradient: sum over the chunk's pairs
      for (int nn = lane; nn < N; nn += 64) {
        float sacc = 0.f;
        for (int p = 0; p < kTrainPairs; ++p) sacc += from_bf16(dz[nn * kTrainPairs + p]);
        atomicAdd(&g_s[boff + nn], sacc);
      }
    };
    wgrad(nullptr, 0, true, &d1_s[0][0], kIn, kHid, kW1, kB1);
    wgrad(&a1_s[0][0], kTrainPairs, false, &d2_s[0][0], kHid, kHid, kW2, kB2);
    wgrad(&a2_s[0][0], kTrainPairs, false, &d3_s[0][0], kHid, kOut, kW3, kB3);
    __syncthreads();
  }
  __syncthreads();
  for (int i = lane; i < kNumParams; i += 64)
    if (g_s[i] != 0.f) atomicAdd(&a.grad[i], g_s[i]);
  for (int o = 32; o > 0; o >>= 1) loss_acc += __shfl_xor(loss_acc, o);
  if (lane == 0) atomicAdd(a.loss, loss_acc * inv_b);
}

// Adam on the FP32 parameters; padded input rows of W1 and unused outputs
// of W3/b3 stay zero.
__global__ void k_adam(float* __restrict__ W, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, float lr, float b1t, float b2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  bool frozen = false;
  if (i < kB1) frozen = (i / kHid) >= kInUsed;
  if (i >= kW3 && i < kB3) frozen = ((i - kW3) % kOut) >= kBaryOut + 4;
  if (i >= kB3) frozen = (i - kB3) >= kBaryOut + 4;
  if (frozen) {
    W[i] = 0.f;
    return;
  }
  const float gi = grad[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
  W[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// ------------------------------------------------------------- inference
// Workgroup of kInferWaves waves; the network's A-operand tiles and biases
// live once per workgroup in LDS, so registers hold only the pair being
// certified and the layer accumulators.  Four waves (one per SIMD) at three
// workgroups per CU: the decision needs up to 168 registers, so three waves
// per SIMD is the occupancy without spills, which an eight-wave workgroup
// (two waves per SIMD each) could not use.
constexpr int kInferWaves = 4;
// A-operand tiles (16 x 32) of the three layers for v_mfma_f32_16x16x32:
// W1 4 output tiles x 1 k-block, W2 4 x 2, W3 (first output tile only) 1 x 2.
constexpr int kTiles = 4 + 8 + 2 + 2;  // + W3 output tile 1 (barycentric outputs)

// Inference modes of k_learned.
enum LearnedMode : int {
  kModeFeatures = 0,  // feature guesses + support check, ties folded (DH_PAIR_LEARNED)
  kModeWarm = 1,      // network barycentrics as the penalty start (DH_PAIR_LEARNED_PENALTY)
  kModeCold = 2       // the same penalty from the reference's 0.33 start (DH_PAIR_PENALTY_CERT)
};
constexpr int kCol = 68;             // row stride (dwords) of the per-wave staging
constexpr int kSubs = 2;             // 16-pair sub-tiles evaluated together

struct LearnedArgs {
  const double* soa;
  int64_t n;
  const float* W;
  double* P;
  double* Q;
  double* d;
  uint8_t* flags;
  // Pairs for the brute force: workgroup b appends to its own region
  // [b * region, (b + 1) * region) of the queue (original index) and of qsoa
  // (coordinates, 18 x qcap, region-major) through a counter in LDS; a
  // single device-wide append counter would serialise every wave on one
  // address across the eight XCDs.
  int* queue;
  double* qsoa;
  int64_t qcap;
  int64_t region;
  unsigned* wg_counts;  // per workgroup: certified, features evaluated, queued
};

// Wait for this wave's LDS operations only (lgkmcnt(0); vmcnt and expcnt at
// their maxima), so the previous chunk's global stores stay in flight.
__device__ inline void lds_wait() { __builtin_amdgcn_s_waitcnt(0xC07F); }

__device__ inline v8s lds_tile(const v8s* tiles, int t, int lane) { return tiles[t * 64 + lane]; }

// Slot (g, j) of k-block b of a hidden layer's input: the previous layer's
// accumulator layout puts units 16t + 4g + j' (j' < 4) of output tile t in
// lane group g, so block b pairs tiles 2b (slots 0..3) and 2b+1 (slots 4..7).
__device__ inline int hidden_k(int b, int g, int j) {
  return 32 * b + 4 * g + (j < 4 ? j : 12 + j);
}

// Tile t of the staged network: lane (r = l%16, g = l/16) holds the 8
// weights W[k][16*to + r] for the k of its slots.  Layer 1's k-block holds
// inputs 8g + j.
__device__ inline v8s stage_tile(const float* W, int t, int lane) {
  const int r = lane & 15, g = lane >> 4;
  float w[8];
  for (int j = 0; j < 8; ++j) {
    if (t < 4)
      w[j] = W[kW1 + (8 * g + j) * kHid + 16 * t + r];
    else if (t < 12)
      w[j] = W[kW2 + hidden_k((t - 4) % 2, g, j) * kHid + 16 * ((t - 4) / 2) + r];
    else if (t < 14)
      w[j] = W[kW3 + hidden_k(t - 12, g, j) * kOut + r];
    else
      w[j] = W[kW3 + hidden_k(t - 14, g, j) * kOut + 16 + r];
  }
  return cat8(pack4(w[0], w[1], w[2], w[3]), pack4(w[4], w[5], w[6], w[7]));
}

// Workgroup LDS of the learned kernels: the network's A-operand tiles and
// biases (staged once per workgroup), and per wave one [16][kCol] dword
// array in column layout (column = pair): first the inputs as 9 rows of bf16
// pairs (inputs 2r, 2r+1), then, for the same columns, the 16 outputs as
// FP32 rows.  A column's inputs are read by the sub-tile that later writes
// its outputs.  kCol = 64 + 4 spreads per-lane rows and the MFMA
// operand/result patterns over all 32 banks (two lanes per bank, the minimum
// for 64 lanes).
struct LearnedLds {
  v8s w_s[kTiles * 64];
  float b_s[2 * kHid + kOut];
  uint32_t io_s[kInferWaves][16 * kCol];
};

// Stage the network once per workgroup (no network in the cold mode: W may
// be null there).  Ends with a workgroup barrier.
template <int MODE>
__device__ __forceinline__ void stage_network(LearnedLds& L, const float* W) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = wave; t < (MODE == kModeCold ? 0 : kTiles); t += kInferWaves)
    L.w_s[t * 64 + lane] = stage_tile(W, t, lane);
  for (int i = threadIdx.x; i < (MODE == kModeCold ? 0 : 2 * kHid + kOut); i += 64 * kInferWaves)
    L.b_s[i] = (i < kHid) ? W[kB1 + i] : (i < 2 * kHid) ? W[kB2 + i - kHid]
                                                        : W[kB3 + i - 2 * kHid];
  __syncthreads();
}

// The learned decision for the wave's 64 pairs (one per lane): canonical
// inputs, network, the two best-ranked features in binary64, the support
// test and the tied features (or, in the penalty modes, the Newton solve
// and its certificate).  done: P, Q are the result; otherwise the pair goes
// to the brute force.  Wave-uniform control flow: every lane of the wave
// must call it.
template <int MODE>
__device__ __forceinline__ void learned_pair(LearnedLds& L, const Tri<double>& A,
                                             const Tri<double>& B, bool valid, V3<double>& P,
                                             V3<double>& Q, bool& done, unsigned& evals) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  v8s* w_s = L.w_s;
  float* b_s = L.b_s;
  uint32_t (*io_s)[16 * kCol] = L.io_s;
  bool frame = true;
  if (MODE != kModeCold) {
    float in[18];
    frame = canonical_inputs(A, B, in);
    uint32_t* col = &io_s[wave][lane];
#pragma unroll
    for (int r = 0; r < 9; ++r)
      col[r * kCol] = static_cast<uint32_t>(bf16_bits(in[2 * r])) |
                      (static_cast<uint32_t>(bf16_bits(in[2 * r + 1])) << 16);
  }
  lds_wait();
  __builtin_amdgcn_wave_barrier();
  // kSubs sub-tiles of 16 pairs at a time, layer by layer: independent
  // MFMA chains share each weight tile read from LDS.
#pragma unroll 1
  for (int s0 = 0; s0 < (MODE == kModeCold ? 0 : 4); s0 += kSubs) {
    // opaque per group: keeps the compiler from hoisting the weight and
    // bias loads out of the loops into ~100 registers
    int wl = lane, bl = 4 * g;
    asm volatile("" : "+v"(wl), "+v"(bl));
    // B operands of layer 1: lane (c, g) holds inputs 8g..8g+7 of its
    // pair (dword rows 4g..4g+3; only inputs 0..17 are nonzero)
    v8s x[kSubs];
#pragma unroll
    for (int u = 0; u < kSubs; ++u) {
      const int pl = (s0 + u) * 16 + c;
      x[u] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (g < 2) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t w = io_s[wave][(4 * g + r) * kCol + pl];
          x[u][2 * r] = static_cast<short>(w & 0xffffu);
          x[u][2 * r + 1] = static_cast<short>(w >> 16);
        }
      } else if (g == 2) {
        const uint32_t w = io_s[wave][8 * kCol + pl];
        x[u][0] = static_cast<short>(w & 0xffffu);
        x[u][1] = static_cast<short>(w >> 16);
      }
    }
    v4s act1[kSubs][4], act2[kSubs][4];
#pragma unroll
    for (int to = 0; to < 4; ++to) {
      const v8s w = lds_tile(w_s, to, wl);
      v4f b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = b_s[to * 16 + bl + j];
#pragma unroll
      for (int u = 0; u < kSubs; ++u) {
        const v4f acc = mfma32(w, x[u], b);
        act1[u][to] = relu4(pack4(acc[0], acc[1], acc[2], acc[3]));
      }
    }
#pragma unroll
    for (int to = 0; to < 4; ++to) {
      const v8s w0 = lds_tile(w_s, 4 + 2 * to, wl), w1 = lds_tile(w_s, 5 + 2 * to, wl);
      v4f b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = b_s[kHid + to * 16 + bl + j];
#pragma unroll
      for (int u = 0; u < kSubs; ++u) {
        v4f acc = mfma32(w0, cat8(act1[u][0], act1[u][1]), b);
        acc = mfma32(w1, cat8(act1[u][2], act1[u][3]), acc);
        act2[u][to] = relu4(pack4(acc[0], acc[1], acc[2], acc[3]));
      }
    }
    if (MODE == kModeFeatures) {  // output tile 0: feature logits and the exact-path logit
      const v8s w0 = lds_tile(w_s, 12, wl), w1 = lds_tile(w_s, 13, wl);
      v4f b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = b_s[2 * kHid + bl + j];
#pragma unroll
      for (int u = 0; u < kSubs; ++u) {
        v4f acc = mfma32(w0, cat8(act2[u][0], act2[u][1]), b);
        acc = mfma32(w1, cat8(act2[u][2], act2[u][3]), acc);
        const int pl = (s0 + u) * 16 + c;
#pragma unroll
        for (int j = 0; j < 4; ++j) io_s[wave][(4 * g + j) * kCol + pl] = __float_as_uint(acc[j]);
      }
    } else {  // output tile 1: outputs 16..19 (lane group 0) = the penalty start, to rows 0..3
      const v8s w0 = lds_tile(w_s, 14, wl), w1 = lds_tile(w_s, 15, wl);
      v4f b;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = b_s[2 * kHid + 16 + bl + j];
#pragma unroll
      for (int u = 0; u < kSubs; ++u) {
        v4f acc = mfma32(w0, cat8(act2[u][0], act2[u][1]), b);
        acc = mfma32(w1, cat8(act2[u][2], act2[u][3]), acc);
        const int pl = (s0 + u) * 16 + c;
        if (g == 0) {
#pragma unroll
          for (int j = 0; j < 4; ++j) io_s[wave][j * kCol + pl] = __float_as_uint(acc[j]);
        }
      }
    }
  }
  lds_wait();
  __builtin_amdgcn_wave_barrier();
  done = false;
  if (MODE == kModeFeatures) {
    // the lane's own pair: the best-ranked edge-edge and the best-ranked
    // vertex-face feature, both evaluated (no divergence between the two
    // kinds), the nearer one checked
    int fe = 0, fv = 9;
    float le = -3.0e38f, lv = -3.0e38f;
#pragma unroll
    for (int f = 0; f < kNumFeat; ++f) {
      const float v = __uint_as_float(io_s[wave][f * kCol + lane]);
      if (f < 9) {
        fe = v > le ? f : fe;
        le = v > le ? v : le;
      } else {
        fv = v > lv ? f : fv;
        lv = v > lv ? v : lv;
      }
    }
    const bool exact_path =
        !frame || (__uint_as_float(io_s[wave][kFlagOut * kCol + lane]) > 0.f);
    V3<double> Pe, Qe, Pv, Qv;
    unsigned sAe, sBe, sAv, sBv;
    double Ve = eval_ee_value(A, B, fe, Pe, Qe, sAe, sBe);
    // one feature at a time: an empty asm makes the vertex-face code depend
    // on the edge-edge result, so the two evaluations are not interleaved
    asm volatile("" : "+v"(fv) : "v"(Ve));
    double Vv = eval_vf_value(A, B, fv, Pv, Qv, sAv, sBv);
    // bfSolver's final rule between its two minima (bf.cpp:1102)
    const bool take_e = Vv > Ve;
    const int win0 = take_e ? fe : fv;
    unsigned open = kUnresolvedAll;
    if (valid && !exact_path)
      open = support_choice(A, B, sel3(take_e, Pe, Pv), sel3(take_e, Qe, Qv),
                            take_e ? sAe : sAv, take_e ? sBe : sBv, win0);
    evals += (valid && !exact_path) ? 2 : 0;
    // Ties of a degenerate closest pair (a vertex against an edge or a
    // vertex): the few features that reach the same points are evaluated
    // and folded into the two minima in bfSolver's order, each minimum
    // keeping its points.  The wave runs its lanes' edge-edge ones
    // together, then its vertex-face ones.
    const unsigned rest = (open == kUnresolvedAll) ? 0u : open & ~((1u << fe) | (1u << fv));
    unsigned rest_e = rest & 0x1FFu, rest_v = rest & 0x7E00u;
    while (__ballot(rest_e != 0u)) {
      if (rest_e) {
        const int g = __builtin_ctz(rest_e);
        rest_e &= rest_e - 1u;
        V3<double> Pg, Qg;
        const double val = eval_ee_value(A, B, g, Pg, Qg);
        const bool better = val < Ve || (val == Ve && g < fe);
        fe = better ? g : fe;
        Ve = better ? val : Ve;
        Pe = sel3(better, Pg, Pe);
        Qe = sel3(better, Qg, Qe);
        ++evals;
      }
    }
    while (__ballot(rest_v != 0u)) {
      if (rest_v) {
        const int g = __builtin_ctz(rest_v);
        rest_v &= rest_v - 1u;
        V3<double> Pg, Qg;
        const double val = eval_vf_value(A, B, g, Pg, Qg);
        const bool better = val < Vv || (val == Vv && g < fv);
        fv = better ? g : fv;
        Vv = better ? val : Vv;
        Pv = sel3(better, Pg, Pv);
        Qv = sel3(better, Qg, Qv);
        ++evals;
      }
    }
    // the final pick (bfSolver's rule again on the folded minima)
    const bool fin_e = Vv > Ve;
    P = sel3(fin_e, Pe, Pv);
    Q = sel3(fin_e, Qe, Qv);
    done = valid && open != kUnresolvedAll;
  } else {
    // penalty Newton from the network's barycentric guess (or 0.33), the
    // result clamped onto the triangles and certified like a feature
    double x0[4] = {0.33, 0.33, 0.33, 0.33};
    if (MODE == kModeWarm)
      for (int k = 0; k < 4; ++k) x0[k] = __uint_as_float(io_s[wave][k * kCol + lane]);
    const int its = warm_penalty(A, B, x0, P, Q);
    done = valid && certify(A, B, P, Q);
    evals += valid ? its : 0;
  }
}

template <int MODE>
__device__ __forceinline__ void learned_body(const LearnedArgs& a) {
  __shared__ LearnedLds L;
  __shared__ unsigned wg_tally[3];  // certified, features evaluated, queued
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x < 3) wg_tally[threadIdx.x] = 0;
  stage_network<MODE>(L, a.W);
  unsigned cert_count = 0, evals = 0;  // per lane
  const int64_t qbase = static_cast<int64_t>(blockIdx.x) * a.region;
  const int64_t nchunks = (a.n + 63) / 64;
  for (int64_t ch = static_cast<int64_t>(blockIdx.x) * kInferWaves + wave; ch < nchunks;
       ch += static_cast<int64_t>(gridDim.x) * kInferWaves) {
    const int64_t i = ch * 64 + lane;
    const bool valid = i < a.n;
    const int64_t ii = valid ? i : 0;
    const Tri<double> A = load_tri_soa(a.soa, a.n, ii, 0);
    const Tri<double> B = load_tri_soa(a.soa, a.n, ii, 9);
    V3<double> P, Q;
    bool done;
    learned_pair<MODE>(L, A, B, valid, P, Q, done, evals);
    // every lane stores (the brute force overwrites the uncertified ones):
    // whole cache lines, no partial-line writes
    if (valid) {
      a.P[i] = P.x;
      a.P[a.n + i] = P.y;
      a.P[2 * a.n + i] = P.z;
      a.Q[i] = Q.x;
      a.Q[a.n + i] = Q.y;
      a.Q[2 * a.n + i] = Q.z;
      a.d[i] = pq_dist(P, Q);
      if (a.flags) a.flags[i] = done ? DH_FLAG_CERTIFIED : DH_FLAG_RECHECK;
      cert_count += done ? 1 : 0;
    }
    // queue the rest for the FP64 brute force: wave-aggregated append of the
    // index and the coordinates (contiguous, so the brute force reads them
    // coalesced)
    const bool need = valid && !done;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(&wg_tally[2], static_cast<unsigned>(__popcll(mask)));
      base = __shfl(base, 0);
      const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
      if (need) {
        const int64_t q = qbase + base + pre;
        a.queue[q] = static_cast<int>(i);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          a.qsoa[(0 + k) * a.qcap + q] = A.v[k].x;
          a.qsoa[(3 + k) * a.qcap + q] = A.v[k].y;
          a.qsoa[(6 + k) * a.qcap + q] = A.v[k].z;
          a.qsoa[(9 + k) * a.qcap + q] = B.v[k].x;
          a.qsoa[(12 + k) * a.qcap + q] = B.v[k].y;
          a.qsoa[(15 + k) * a.qcap + q] = B.v[k].z;
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cert_count += __shfl_xor(cert_count, o);
    evals += __shfl_xor(evals, o);
  }
  if (lane == 0) {
    atomicAdd(&wg_tally[0], cert_count);
    atomicAdd(&wg_tally[1], evals);
  }
  __syncthreads();
  if (threadIdx.x < 3) a.wg_counts[3 * blockIdx.x + threadIdx.x] = wg_tally[threadIdx.x];
  // Drain the workgroup's queue with the FP64 brute force, every lane busy:
  // the entries were written by this workgroup (same CU, L2-resident), and
  // their placeholder results are overwritten here.
  const unsigned qlen = wg_tally[2];
  for (unsigned e0 = wave * 64; e0 < qlen; e0 += kInferWaves * 64) {
    const unsigned e = e0 + lane;
    if (e < qlen) {
      const int64_t q = qbase + e;
      const int64_t i = a.queue[q];
      const Tri<double> A = load_tri_soa(a.qsoa, a.qcap, q, 0);
      const Tri<double> B = load_tri_soa(a.qsoa, a.qcap, q, 9);
      V3<double> p, qq;
      const bool edge = bf_solver(A, B, p, qq);
      a.P[i] = p.x;
      a.P[a.n + i] = p.y;
      a.P[2 * a.n + i] = p.z;
      a.Q[i] = qq.x;
      a.Q[a.n + i] = qq.y;
      a.Q[2 * a.n + i] = qq.z;
      a.d[i] = pq_dist(p, qq);
      if (a.flags) a.flags[i] = DH_FLAG_RECHECK | (edge ? DH_FLAG_EDGE : 0);
    }
  }
}

// Three waves per SIMD (168 registers; the tie folding keeps both families'
// minima and their points live, and at four waves the registers spill).
template <int MODE>
__global__ __launch_bounds__(64 * kInferWaves) __attribute__((amdgpu_waves_per_eu(3))) void
k_learned(LearnedArgs a) {
  learned_body<MODE>(a);
}

// ------------------------------------------------ the step's survivor queue
// The triangle pairs that the DEM step's box mid-phase leaves (the survivor
// queue of detect.hip, entries {item, iA | iB << 16}) through the learned
// decision: a wave per 64 entries on a persistent grid.  A finished entry
// sets flags[e] (its pair emits a contact: d <= epsA + epsB, bf.cpp:125) and,
// when it emits, its P, Q at pq[6e..6e+6); the rest are listed for the
// brute force (k_queue_fallback).  Both write exactly what the brute-force
// queue solve would.
struct QueueLearnArgs {
  const int2* queue;
  const int* pairs;
  int tiles;  // items per particle pair
  const int* tri_offset;
  const double* x;
  const double* y;
  const double* z;
  const double* eps;
  long long n;                         // entries of the launch
  const unsigned long long* need_dev;  // live entries (captured step), else null
  int* flags;
  double* pq;
  const float* W;
  int* fb_list;
  unsigned* fb_count;
};

__device__ __forceinline__ Tri<double> load_tri_xyz(const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z, int t) {
  Tri<double> r;
#pragma unroll
  for (int k = 0; k < 3; ++k) r.v[k] = {x[3 * t + k], y[3 * t + k], z[3 * t + k]};
  return r;
}

__global__ __launch_bounds__(64 * kInferWaves) __attribute__((amdgpu_waves_per_eu(3))) void
k_queue_learned(QueueLearnArgs a) {
  __shared__ LearnedLds L;
  stage_network<kModeFeatures>(L, a.W);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long live =
      a.need_dev && static_cast<long long>(*a.need_dev) < a.n ? static_cast<long long>(*a.need_dev)
                                                              : a.n;
  // the live entries only (the later passes stop at the same count)
  const long long nchunks = (live + 63) / 64;
  unsigned evals = 0;
  for (long long ch = static_cast<long long>(blockIdx.x) * kInferWaves + wave; ch < nchunks;
       ch += static_cast<long long>(gridDim.x) * kInferWaves) {
    const long long e = ch * 64 + lane;
    const int2 en = e < live ? a.queue[e] : make_int2(-1, 0);
    const bool valid = en.x >= 0;
    Tri<double> A = {}, B = {};
    double margin = 0.0;
    if (valid) {
      const int k = en.x / a.tiles;
      const int pa = a.pairs[2 * k], pb = a.pairs[2 * k + 1];
      const int iA = static_cast<int>(static_cast<unsigned>(en.y) & 0xffffu);
      const int iB = static_cast<int>(static_cast<unsigned>(en.y) >> 16);
      A = load_tri_xyz(a.x, a.y, a.z, a.tri_offset[pa] + iA);
      B = load_tri_xyz(a.x, a.y, a.z, a.tri_offset[pb] + iB);
      margin = a.eps[pa] + a.eps[pb];
    }
    V3<double> P, Q;
    bool done;
    learned_pair<kModeFeatures>(L, A, B, valid, P, Q, done, evals);
    if (e < live && !(valid && !done)) {
      const bool emit = valid && pq_dist(P, Q) <= margin;
      a.flags[e] = emit ? 1 : 0;
      if (emit) {
        double* o = a.pq + 6 * e;
        o[0] = P.x; o[1] = P.y; o[2] = P.z;
        o[3] = Q.x; o[4] = Q.y; o[5] = Q.z;
      }
    }
    // the rest to the brute force: one append per wave
    const bool need = valid && !done;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(a.fb_count, static_cast<unsigned>(__popcll(mask)));
      base = __shfl(base, 0);
      const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
      if (need) a.fb_list[base + pre] = static_cast<int>(e);
    }
  }
}

__device__ __forceinline__ void queue_fallback_entry(const QueueLearnArgs& a, long long e) {
  const int2 en = a.queue[e];
  const int k = en.x / a.tiles;
  const int pa = a.pairs[2 * k], pb = a.pairs[2 * k + 1];
  const int iA = static_cast<int>(static_cast<unsigned>(en.y) & 0xffffu);
  const int iB = static_cast<int>(static_cast<unsigned>(en.y) >> 16);
  const Tri<double> A = load_tri_xyz(a.x, a.y, a.z, a.tri_offset[pa] + iA);
  const Tri<double> B = load_tri_xyz(a.x, a.y, a.z, a.tri_offset[pb] + iB);
  V3<double> P, Q;
  bf_solver(A, B, P, Q);
  const bool emit = pq_dist(P, Q) <= a.eps[pa] + a.eps[pb];
  a.flags[e] = emit ? 1 : 0;
  if (emit) {
    double* o = a.pq + 6 * e;
    o[0] = P.x; o[1] = P.y; o[2] = P.z;
    o[3] = Q.x; o[4] = Q.y; o[5] = Q.z;
  }
}

// The listed entries by the brute force, a lane each at a time (a captured
// step launches a persistent grid and strides over the list's length).
__global__ __launch_bounds__(256) void k_queue_fallback(QueueLearnArgs a) {
  const long long nfb = static_cast<long long>(*a.fb_count) < a.n
                            ? static_cast<long long>(*a.fb_count) : a.n;
  for (long long j = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; j < nfb;
       j += static_cast<long long>(gridDim.x) * 256)
    queue_fallback_entry(a, a.fb_list[j]);
}

struct LearnedState {
  DevBuf queue, qsoa, wg_counts, model_copy;
  const float* model = nullptr;  // points into model_copy once set
  unsigned grid = 0;              // workgroups of the last launch
};
static LearnedState g_learned;

struct QueueLearnState {
  DevBuf fb_list, fb_count;
};
static QueueLearnState g_qlearn;

// Buffers of the learned queue solve for n entries (before a capture).
int queue_learned_reserve(long long n) {
  int rc;
  if ((rc = g_qlearn.fb_list.grow(sizeof(int) * (n + 1)))) return rc;
  return g_qlearn.fb_count.grow(sizeof(unsigned));
}

// The survivor queue's n entries (live: *need_dev when set) through the
// learned decision, then the brute force for the rest; fills instead of
// memsets (captured steps).
int queue_learned(const int2* queue, const int* pairs, int tiles, const int* tri_offset,
                  const double* x, const double* y, const double* z, const double* eps,
                  long long n, const unsigned long long* need_dev, int* flags, double* pq,
                  hipStream_t st) {
  if (!g_learned.model)
    return set_error(DH_ERR_ARG, "DH_BF_LEARNED: no network set (dh_learned_set_model)");
  int rc;
  if ((rc = queue_learned_reserve(n))) return rc;
  QueueLearnArgs a{queue, pairs, tiles, tri_offset, x, y, z, eps, n, need_dev, flags, pq,
                   g_learned.model, g_qlearn.fb_list.as<int>(), g_qlearn.fb_count.as<unsigned>()};
  fill_async(g_qlearn.fb_count.as<unsigned>(), 0u, 1, st);
  if (n > 0) {
    const long long chunks = (n + 63) / 64;
    long long grid = (chunks + kInferWaves - 1) / kInferWaves;
    const long long max_grid = static_cast<long long>(device_cus()) * 4;
    if (grid > max_grid) grid = max_grid;
    hipLaunchKernelGGL(k_queue_learned, static_cast<unsigned>(grid), 64 * kInferWaves, 0, st, a);
    long long fgrid = grid_for(n, 256);
    const long long fcap = static_cast<long long>(device_cus()) * 16;
    if (need_dev && fgrid > fcap) fgrid = fcap;
    hipLaunchKernelGGL(k_queue_fallback, static_cast<unsigned>(fgrid), 256, 0, st, a);
  }
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int learned_pairs(int mode, const double* soa, int64_t n, double* P, double* Q, double* d,
                  uint8_t* flags, hipStream_t st) {
  if (mode != kModeCold && !g_learned.model)
    return set_error(DH_ERR_ARG, "learned narrow phase: no network set (dh_learned_set_model)");
  const int64_t chunks = (n + 63) / 64;
  int64_t grid = (chunks + kInferWaves - 1) / kInferWaves;
  // persistent: at most learned_wg_per_cu workgroups per CU
  const int64_t max_grid = static_cast<int64_t>(device_cus()) * tuning().learned_wg_per_cu;
  if (grid > max_grid) grid = max_grid;
  if (grid < 1) grid = 1;
  // every wave of a workgroup takes at most `per_wave` chunks
  const int64_t per_wave = (chunks + grid * kInferWaves - 1) / (grid * kInferWaves);
  const int64_t region = per_wave * kInferWaves * 64;
  const int64_t qcap = grid * region;
  int rc;
  if ((rc = g_learned.queue.grow(sizeof(int) * qcap))) return rc;
  if ((rc = g_learned.qsoa.grow(sizeof(double) * 18 * qcap))) return rc;
  if ((rc = g_learned.wg_counts.grow(sizeof(unsigned) * 3 * grid))) return rc;
  g_learned.grid = static_cast<unsigned>(grid);
  LearnedArgs a{soa, n, g_learned.model, P, Q, d, flags, g_learned.queue.as<int>(),
                g_learned.qsoa.as<double>(), qcap, region, g_learned.wg_counts.as<unsigned>()};
  if (n > 0) {
    if (mode == kModeWarm)
      hipLaunchKernelGGL(k_learned<kModeWarm>, static_cast<unsigned>(grid), 64 * kInferWaves,
                         0, st, a);
    else if (mode == kModeCold)
      hipLaunchKernelGGL(k_learned<kModeCold>, static_cast<unsigned>(grid), 64 * kInferWaves,
                         0, st, a);
    else
      hipLaunchKernelGGL(k_learned<kModeFeatures>, static_cast<unsigned>(grid), 64 * kInferWaves,
                         0, st, a);
  }
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

// Totals of the last learned pass: certified, features evaluated, queued.
int learned_counters(unsigned long long out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (!g_learned.wg_counts.ptr || g_learned.grid == 0) return DH_OK;
  std::vector<unsigned> h(3 * static_cast<size_t>(g_learned.grid));
  DH_HIP_CHECK(hipMemcpy(h.data(), g_learned.wg_counts.ptr, sizeof(unsigned) * h.size(),
                         hipMemcpyDeviceToHost));
  for (size_t b = 0; b < g_learned.grid; ++b)
    for (int k = 0; k < 3; ++k) out[k] += h[3 * b + k];
  return DH_OK;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_mlp_param_count(void) { return kNumParams; }

int dh_mlp_labels(const double* soa_dev, int64_t n, uint16_t* X_dev, uint8_t* feat_dev,
                  float* bary_dev, float* flag_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!soa_dev || !X_dev || !feat_dev || !bary_dev || !flag_dev)))
    return set_error(DH_ERR_ARG, "dh_mlp_labels: bad arguments");
  if (n == 0) return DH_OK;
  hipLaunchKernelGGL(k_labels, grid_for(n, 256), 256, 0, as_stream(stream), soa_dev, n, X_dev,
                     feat_dev, bary_dev, flag_dev);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_mlp_forward(const float* params_dev, const uint16_t* X_dev, int64_t n, float* out_dev,
                   void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!params_dev || !X_dev || !out_dev)))
    return set_error(DH_ERR_ARG, "dh_mlp_forward: bad arguments");
  if (n == 0) return DH_OK;
  int64_t tiles = (n + 15) / 16;
  if (tiles > 65536) tiles = 65536;
  hipLaunchKernelGGL(k_mlp_forward, static_cast<unsigned>(tiles), 64, 0, as_stream(stream),
                     params_dev, X_dev, n, out_dev);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_mlp_grad(const float* params_dev, const uint16_t* X_dev, const uint8_t* feat_dev,
                const float* bary_dev, const float* flag_dev, int64_t offset, int batch,
                float* grad_dev, float* loss_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (batch <= 0 || !params_dev || !X_dev || !feat_dev || !bary_dev || !flag_dev || !grad_dev ||
      !loss_dev)
    return set_error(DH_ERR_ARG, "dh_mlp_grad: bad arguments");
  const hipStream_t st = as_stream(stream);
  DH_HIP_CHECK(hipMemsetAsync(grad_dev, 0, sizeof(float) * kNumParams, st));
  DH_HIP_CHECK(hipMemsetAsync(loss_dev, 0, sizeof(float), st));
  TrainArgs a{params_dev, X_dev, feat_dev, bary_dev, flag_dev, offset, batch, grad_dev, loss_dev};
  int grid = (batch + kTrainPairs - 1) / kTrainPairs;
  if (grid > 1024) grid = 1024;
  hipLaunchKernelGGL(k_mlp_train, grid, 64, 0, st, a);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_mlp_adam(float* params_dev, float* m_dev, float* v_dev, const float* grad_dev, float lr,
                int step, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!params_dev || !m_dev || !v_dev || !grad_dev || step < 1)
    return set_error(DH_ERR_ARG, "dh_mlp_adam: bad arguments");
  const float b1t = powf(0.9f, static_cast<float>(step));
  const float b2t = powf(0.999f, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, grid_for(kNumParams, 256), 256, 0, as_stream(stream), params_dev,
                     m_dev, v_dev, grad_dev, lr, b1t, b2t);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_learned_set_model(const float* params_dev) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!params_dev) return set_error(DH_ERR_ARG, "dh_learned_set_model: null");
  if ((rc = g_learned.model_copy.grow(sizeof(float) * kNumParams))) return rc;
  DH_HIP_CHECK(hipMemcpy(g_learned.model_copy.ptr, params_dev, sizeof(float) * kNumParams,
                         hipMemcpyDeviceToDevice));
  g_learned.model = g_learned.model_copy.as<float>();
  return DH_OK;
}

}  // extern "C"

namespace dh {
int tri_pairs_learned(int method, const double* soa, int64_t n, double* P, double* Q,
                      double* d, uint8_t* flags, hipStream_t st) {
  const int mode = method == DH_PAIR_LEARNED_PENALTY ? kModeWarm
                   : method == DH_PAIR_PENALTY_CERT  ? kModeCold
                                                     : kModeFeatures;
  return learned_pairs(mode, soa, n, P, Q, d, flags, st);
}
}  // namespace dh
