// learned.hip — the learned narrow phase: labels from the FP64 brute force,
// MFMA training of the feature network, and the fused inference kernel.
//
// Inference per pair: canonical-frame inputs -> network (bf16 MFMA) -> the two
// most likely closest features -> each evaluated in binary64 with bfSolver's
// arithmetic and checked by the optimality certificate (pair_features.hpp).
// Certified pairs are finished; the others (and those the network marks for
// the exact path) are compacted into a queue that the FP64 brute-force kernel
// drains.  The network therefore decides only how much work a pair costs;
// every distance is either certified within 8e-7 relative or bfSolver's own.
#include "exact_math.hpp"
#include "mlp.hpp"
#include "pair_features.hpp"
#include "runtime.hpp"

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
  const int code = bf_feature_code(A, B, P, Q);
  feat[i] = static_cast<uint8_t>(code);
  double al, be, ga, de;
  barycentric(A, P, al, be);
  barycentric(B, Q, ga, de);
  auto clampf = [](double v) { return static_cast<float>(v < -0.5 ? -0.5 : (v > 1.5 ? 1.5 : v)); };
  bary[4 * i + 0] = clampf(al);
  bary[4 * i + 1] = clampf(be);
  bary[4 * i + 2] = clampf(ga);
  bary[4 * i + 3] = clampf(de);
  flag[i] = (frame && certify(A, B, P, Q)) ? 0.0f : 1.0f;
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
    a1[to] = pack4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
  }
#pragma unroll
  for (int to = 0; to < 4; ++to) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) acc = mfma16(fw.w2[to][tk], a1[tk], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] += W[kB2 + to * 16 + 4 * g + j];
    z2[to] = acc;
    a2[to] = pack4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
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
          a1_s[to * 16 + 4 * g + j][pl] = bf16_bits(gelu(z1[to][j]));
          a2_s[to * 16 + 4 * g + j][pl] = bf16_bits(gelu(z2[to][j]));
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
          d[j] = acc[j] * gelu_grad(z2[tu][j]);
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
          d1_s[tu * 16 + 4 * g + j][pl] = bf16_bits(acc[j] * gelu_grad(z1[tu][j]));
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
constexpr int kInferWaves = 4;

struct LearnedArgs {
  const double* soa;
  int64_t n;
  const float* W;
  double* P;
  double* Q;
  double* d;
  uint8_t* flags;
  int* queue;               // indices for the FP64 brute force
  unsigned long long* qlen;
  unsigned long long* counters;  // [0] certified, [1] features evaluated
};

__global__ __launch_bounds__(64 * kInferWaves) void k_learned(LearnedArgs a) {
  __shared__ uint16_t x_s[kInferWaves][64][kIn];
  __shared__ float o_s[kInferWaves][64][16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  FwdWeights fw;
  load_fwd(a.W, fw, lane);
  unsigned long long cert_count = 0, evals = 0;
  const int64_t nchunks = (a.n + 63) / 64;
  for (int64_t ch = static_cast<int64_t>(blockIdx.x) * kInferWaves + wave; ch < nchunks;
       ch += static_cast<int64_t>(gridDim.x) * kInferWaves) {
    const int64_t i = ch * 64 + lane;
    const bool valid = i < a.n;
    const int64_t ii = valid ? i : 0;
    bool frame;
    {
      const Tri<double> A = load_tri_soa(a.soa, a.n, ii, 0);
      const Tri<double> B = load_tri_soa(a.soa, a.n, ii, 9);
      float in[18];
      frame = canonical_inputs(A, B, in);
      for (int k = 0; k < kIn; ++k) x_s[wave][lane][k] = (k < kInUsed) ? bf16_bits(in[k]) : 0;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
      const int pl = s * 16 + c;
      v4s xb[2] = {x_operand(&x_s[wave][pl][0], 0, lane), x_operand(&x_s[wave][pl][0], 1, lane)};
      v4f z1[4], z2[4], z3[2];
      forward16(fw, a.W, xb, lane, z1, z2, z3);
#pragma unroll
      for (int j = 0; j < 4; ++j) o_s[wave][pl][4 * g + j] = z3[0][j];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // the lane's own pair: two best features, then the certificate
    int f1 = 0, f2 = 1;
    float l1 = -3.0e38f, l2 = -3.0e38f;
    for (int f = 0; f < kNumFeat; ++f) {
      const float v = o_s[wave][lane][f];
      if (v > l1) {
        l2 = l1;
        f2 = f1;
        l1 = v;
        f1 = f;
      } else if (v > l2) {
        l2 = v;
        f2 = f;
      }
    }
    const bool exact_path = !frame || (o_s[wave][lane][kFlagOut] > 0.f);
    // the pair again (cache-resident): keeps it out of registers during the network
    const Tri<double> A = load_tri_soa(a.soa, a.n, ii, 0);
    const Tri<double> B = load_tri_soa(a.soa, a.n, ii, 9);
    V3<double> P = {0, 0, 0}, Q = {0, 0, 0};
    bool done = false;
    if (valid && !exact_path) {
      eval_feature(A, B, f1, P, Q);
      evals++;
      done = certify(A, B, P, Q);
    }
    if (__ballot(valid && !exact_path && !done)) {
      if (valid && !exact_path && !done) {
        eval_feature(A, B, f2, P, Q);
        evals++;
        done = certify(A, B, P, Q);
      }
    }
    if (valid && done) {
      a.P[i] = P.x;
      a.P[a.n + i] = P.y;
      a.P[2 * a.n + i] = P.z;
      a.Q[i] = Q.x;
      a.Q[a.n + i] = Q.y;
      a.Q[2 * a.n + i] = Q.z;
      a.d[i] = pq_dist(P, Q);
      if (a.flags) a.flags[i] = DH_FLAG_CERTIFIED;
      cert_count++;
    }
    // queue the rest for the FP64 brute force (wave-aggregated append)
    const bool need = valid && !done;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(a.qlen, static_cast<unsigned long long>(__popcll(mask)));
      base = __shfl(base, 0);
      const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
      if (need) a.queue[base + pre] = static_cast<int>(i);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cert_count += __shfl_xor(cert_count, o);
    evals += __shfl_xor(evals, o);
  }
  if (lane == 0) {
    atomicAdd(&a.counters[0], cert_count);
    atomicAdd(&a.counters[1], evals);
  }
}

// FP64 brute force over the queued pairs (full lane occupancy).
__global__ __launch_bounds__(256) void k_queue_bf(const double* __restrict__ soa, int64_t n,
                                                  const int* __restrict__ queue,
                                                  const unsigned long long* __restrict__ qlen,
                                                  double* __restrict__ P, double* __restrict__ Q,
                                                  double* __restrict__ d,
                                                  uint8_t* __restrict__ flags) {
  const unsigned long long len = *qlen;
  for (unsigned long long t = static_cast<unsigned long long>(blockIdx.x) * 256 + threadIdx.x;
       t < len; t += static_cast<unsigned long long>(gridDim.x) * 256) {
    const int64_t i = queue[t];
    const Tri<double> A = load_tri_soa(soa, n, i, 0), B = load_tri_soa(soa, n, i, 9);
    V3<double> p, q;
    const bool edge = bf_solver(A, B, p, q);
    P[i] = p.x;
    P[n + i] = p.y;
    P[2 * n + i] = p.z;
    Q[i] = q.x;
    Q[n + i] = q.y;
    Q[2 * n + i] = q.z;
    d[i] = pq_dist(p, q);
    if (flags) flags[i] = DH_FLAG_RECHECK | (edge ? DH_FLAG_EDGE : 0);
  }
}

struct LearnedState {
  DevBuf queue, qlen, counters, model_copy;
  const float* model = nullptr;  // points into model_copy once set
  unsigned long long last_counters[3] = {0, 0, 0};
};
static LearnedState g_learned;

int learned_pairs(const double* soa, int64_t n, double* P, double* Q, double* d, uint8_t* flags,
                  hipStream_t st) {
  if (!g_learned.model)
    return set_error(DH_ERR_ARG, "learned narrow phase: no network set (dh_learned_set_model)");
  int rc;
  if ((rc = g_learned.queue.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_learned.qlen.grow(sizeof(unsigned long long)))) return rc;
  if ((rc = g_learned.counters.grow(sizeof(unsigned long long) * 4))) return rc;
  DH_HIP_CHECK(hipMemsetAsync(g_learned.qlen.ptr, 0, sizeof(unsigned long long), st));
  DH_HIP_CHECK(hipMemsetAsync(g_learned.counters.ptr, 0, sizeof(unsigned long long) * 4, st));
  LearnedArgs a{soa, n, g_learned.model, P, Q, d, flags, g_learned.queue.as<int>(),
                g_learned.qlen.as<unsigned long long>(),
                g_learned.counters.as<unsigned long long>()};
  const int64_t chunks = (n + 63) / 64;
  int64_t grid = (chunks + kInferWaves - 1) / kInferWaves;
  if (grid > 256 * 16) grid = 256 * 16;  // persistent: 16 workgroups per CU
  hipLaunchKernelGGL(k_learned, static_cast<unsigned>(grid), 64 * kInferWaves, 0, st, a);
  int64_t qgrid = (n + 255) / 256;
  if (qgrid > 256 * 16) qgrid = 256 * 16;
  hipLaunchKernelGGL(k_queue_bf, static_cast<unsigned>(qgrid), 256, 0, st, soa, n,
                     g_learned.queue.as<int>(), g_learned.qlen.as<unsigned long long>(), P, Q, d,
                     flags);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int learned_counters(unsigned long long out[3]) {
  unsigned long long h[2] = {0, 0}, q = 0;
  if (g_learned.counters.ptr) {
    DH_HIP_CHECK(hipMemcpy(h, g_learned.counters.ptr, sizeof(h), hipMemcpyDeviceToHost));
    DH_HIP_CHECK(hipMemcpy(&q, g_learned.qlen.ptr, sizeof(q), hipMemcpyDeviceToHost));
  }
  out[0] = h[0];
  out[1] = h[1];
  out[2] = q;
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
int tri_pairs_learned(const double* soa, int64_t n, double* P, double* Q, double* d,
                      uint8_t* flags, hipStream_t st) {
  return learned_pairs(soa, n, P, Q, d, flags, st);
}
}  // namespace dh
