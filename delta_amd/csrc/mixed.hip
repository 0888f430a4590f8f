// mixed.hip — the FP32 screen with FP64 recheck for unique triangle pairs, and
// dispatch of the non-exact unique-pair paths.
//
// DH_PAIR_MIXED: every pair is solved by bfSolver's arithmetic in binary32 on
// the caller's binary32 copy (72 bytes per pair), with coordinates taken
// relative to A's vertex 0.  The binary32 solve only picks the closest
// feature pair; the distance is bounded from both sides in binary64 on the
// binary32 triangles (which binary64 holds exactly):
//   * UB = |Q - P| with P, Q that feature's closest points evaluated in
//     binary64 (points on the triangles, up to binary64 rounding);
//   * LB = min over B's vertices of m.b - max over A's vertices of m.a,
//     m = (Q - P)/|Q - P| (a separating direction; its vertices suffice by
//     convexity), less 2^-49 (RA + RB) for the binary64 rounding of the six
//     dot products (RA, RB the largest 1-norms of the relative vertices);
//   * the binary32 copy moved each vertex by at most
//     2^-24 (|v|_1 + |o|_1 + |v - o|_1) (the copy's rounding of v and of A's
//     vertex o, and the rounding of the difference), so the true distance
//     lies within delta (the largest such shift of A plus that of B) of the
//     binary32 triangles' distance.
// Then |UB - d| <= err = (UB - LB) + delta for the true distance d, and
// d >= LB - delta.  A pair is finished in binary32 when LB - delta is at or
// above the recheck band (2 * eps_total) and err <= 1e-6 (LB - delta): its
// distance is within 1e-6 relative of the binary64 brute force's by
// construction, and it is no contact.  Every other pair -- within the band,
// near-parallel or degenerate (the binary32 feature is then not a closest
// pair and UB - LB exceeds the allowance), not finite -- is appended to the
// workgroup's own region of a queue through an LDS counter and, after the
// workgroup's last chunk, recomputed by the binary64 brute force from the
// binary64 coordinates.  Every contact (d <= eps_total) is therefore decided
// in binary64.
#include "exact_math.hpp"
#include "pair_features.hpp"
#include "runtime.hpp"

#include <vector>

namespace dh {

int tri_pairs_learned(int method, const double* soa, int64_t n, double* P, double* Q,
                      double* d, uint8_t* flags, hipStream_t st);
int learned_counters(unsigned long long out[3]);

constexpr int kMixedWaves = 4;

struct MixedArgs {
  const double* soa;
  const float* soa32;
  int64_t n;
  float band;  // recheck when the binary32 distance is below this
  double* P;
  double* Q;
  double* d;
  uint8_t* flags;
  int* queue;           // per-workgroup regions of original pair indices
  int64_t region;
  unsigned* wg_counts;  // per workgroup: finished in binary32, rechecked
};

__device__ __forceinline__ Tri<double> load_tri64(const double* __restrict__ soa, int64_t n,
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

__global__ __launch_bounds__(64 * kMixedWaves) void k_mixed(MixedArgs a) {
  __shared__ unsigned tally[2];
  if (threadIdx.x < 2) tally[threadIdx.x] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t qbase = static_cast<int64_t>(blockIdx.x) * a.region;
  const int64_t nchunks = (a.n + 63) / 64;
  unsigned finished = 0;
  for (int64_t ch = static_cast<int64_t>(blockIdx.x) * kMixedWaves + wave; ch < nchunks;
       ch += static_cast<int64_t>(gridDim.x) * kMixedWaves) {
    const int64_t i = ch * 64 + lane;
    const bool valid = i < a.n;
    const int64_t ii = valid ? i : 0;
    const float* s = a.soa32;
    const int64_t n = a.n;
    const V3<float> o = {s[0 * n + ii], s[3 * n + ii], s[6 * n + ii]};
    Tri<float> A, B;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      A.v[k] = {s[(0 + k) * n + ii] - o.x, s[(3 + k) * n + ii] - o.y, s[(6 + k) * n + ii] - o.z};
      B.v[k] = {s[(9 + k) * n + ii] - o.x, s[(12 + k) * n + ii] - o.y,
                s[(15 + k) * n + ii] - o.z};
    }
    V3<float> p, q;
    int ssid = 0, vid = 0;
    const bool edge = bf_solver_ids(A, B, p, q, ssid, vid);
    // the feature in binary64 on the binary32 triangles (exact in binary64)
    Tri<double> Ad, Bd;
    double RA = 0.0, RB = 0.0, shA = 0.0, shB = 0.0;
    const double o1 = fabs(static_cast<double>(o.x)) + fabs(static_cast<double>(o.y)) +
                      fabs(static_cast<double>(o.z));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Ad.v[k] = {A.v[k].x, A.v[k].y, A.v[k].z};
      Bd.v[k] = {B.v[k].x, B.v[k].y, B.v[k].z};
      const double ra = fabs(Ad.v[k].x) + fabs(Ad.v[k].y) + fabs(Ad.v[k].z);
      const double rb = fabs(Bd.v[k].x) + fabs(Bd.v[k].y) + fabs(Bd.v[k].z);
      RA = fmax(RA, ra);
      RB = fmax(RB, rb);
      // |v|_1 of the absolute vertex <= |v - o|_1 + |o|_1
      shA = fmax(shA, 2.0 * (ra + o1));
      shB = fmax(shB, 2.0 * (rb + o1));
    }
    V3<double> Pd, Qd;
    eval_feature(Ad, Bd, edge ? ssid : 9 + vid, Pd, Qd);
    const V3<double> nv = sub(Qd, Pd);
    const double ub = sqrt(dot(nv, nv)) * (1.0 + 0x1.0p-50);
    const double inv = 1.0 / ub;
    const V3<double> m = {nv.x * inv, nv.y * inv, nv.z * inv};
    double amax = -INFINITY, bmin = INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      amax = fmax(amax, dot(m, Ad.v[k]));
      bmin = fmin(bmin, dot(m, Bd.v[k]));
    }
    const double lb = bmin - amax - 0x1.0p-49 * (RA + RB);
    const double delta = 0x1.0p-24 * (1.0 + 0x1.0p-20) * (shA + shB);
    const double dlo = lb - delta;
    const double err = (ub - lb) + delta;
    const bool keep = valid && dlo >= static_cast<double>(a.band) && err <= 1e-6 * dlo;
    if (keep) {
      const double ox = o.x, oy = o.y, oz = o.z;
      a.P[i] = ox + Pd.x;
      a.P[n + i] = oy + Pd.y;
      a.P[2 * n + i] = oz + Pd.z;
      a.Q[i] = ox + Qd.x;
      a.Q[n + i] = oy + Qd.y;
      a.Q[2 * n + i] = oz + Qd.z;
      a.d[i] = ub;
      if (a.flags) a.flags[i] = edge ? DH_FLAG_EDGE : 0;
      finished++;
    }
    const bool need = valid && !keep;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(&tally[1], static_cast<unsigned>(__popcll(mask)));
      base = __shfl(base, 0);
      const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
      if (need) a.queue[qbase + base + pre] = static_cast<int>(i);
    }
  }
  for (int o = 32; o > 0; o >>= 1) finished += __shfl_xor(finished, o);
  if (lane == 0) atomicAdd(&tally[0], finished);
  __syncthreads();
  if (threadIdx.x < 2) a.wg_counts[2 * blockIdx.x + threadIdx.x] = tally[threadIdx.x];
  // binary64 recheck of the workgroup's queue, every lane busy
  const unsigned qlen = tally[1];
  for (unsigned e0 = wave * 64; e0 < qlen; e0 += kMixedWaves * 64) {
    const unsigned e = e0 + lane;
    if (e < qlen) {
      const int64_t i = a.queue[qbase + e];
      const Tri<double> A = load_tri64(a.soa, a.n, i, 0);
      const Tri<double> B = load_tri64(a.soa, a.n, i, 9);
      V3<double> p, q;
      const bool edge = bf_solver(A, B, p, q);
      a.P[i] = p.x;
      a.P[a.n + i] = p.y;
      a.P[2 * a.n + i] = p.z;
      a.Q[i] = q.x;
      a.Q[a.n + i] = q.y;
      a.Q[2 * a.n + i] = q.z;
      a.d[i] = pq_dist(p, q);
      if (a.flags) a.flags[i] = DH_FLAG_RECHECK | (edge ? DH_FLAG_EDGE : 0);
    }
  }
}

struct MixedState {
  DevBuf queue, wg_counts;
  unsigned grid = 0;
};
static MixedState g_mixed;
static int g_last_extra = -1;  // method of the last non-exact pass

static int tri_pairs_mixed(const double* soa, const float* soa32, int64_t n, double eps_total,
                           double* P, double* Q, double* d, uint8_t* flags, hipStream_t st) {
  if (!soa32)
    return set_error(DH_ERR_ARG, "DH_PAIR_MIXED needs the binary32 copy (soa32_dev, "
                                 "dh_convert_f32)");
  const int64_t chunks = (n + 63) / 64;
  int64_t grid = (chunks + kMixedWaves - 1) / kMixedWaves;
  const int64_t max_grid = static_cast<int64_t>(device_cus()) * tuning().mixed_wg_per_cu;
  if (grid > max_grid) grid = max_grid;
  if (grid < 1) grid = 1;
  const int64_t per_wave = (chunks + grid * kMixedWaves - 1) / (grid * kMixedWaves);
  const int64_t region = per_wave * kMixedWaves * 64;
  int rc;
  if ((rc = g_mixed.queue.grow(sizeof(int) * grid * region))) return rc;
  if ((rc = g_mixed.wg_counts.grow(sizeof(unsigned) * 2 * grid))) return rc;
  g_mixed.grid = static_cast<unsigned>(grid);
  MixedArgs a{soa, soa32, n, static_cast<float>(2.0 * eps_total), P, Q, d, flags,
              g_mixed.queue.as<int>(), region, g_mixed.wg_counts.as<unsigned>()};
  if (n > 0) hipLaunchKernelGGL(k_mixed, static_cast<unsigned>(grid), 64 * kMixedWaves, 0, st, a);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

static int mixed_counters(unsigned long long out[2]) {
  out[0] = out[1] = 0;
  if (!g_mixed.wg_counts.ptr || g_mixed.grid == 0) return DH_OK;
  std::vector<unsigned> h(2 * static_cast<size_t>(g_mixed.grid));
  DH_HIP_CHECK(hipMemcpy(h.data(), g_mixed.wg_counts.ptr, sizeof(unsigned) * h.size(),
                         hipMemcpyDeviceToHost));
  for (size_t b = 0; b < g_mixed.grid; ++b) {
    out[0] += h[2 * b];
    out[1] += h[2 * b + 1];
  }
  return DH_OK;
}

int tri_pairs_extra(int method, const double* soa, const float* soa32, int64_t n,
                    double eps_total, double* P, double* Q, double* d, uint8_t* flags,
                    hipStream_t st) {
  g_last_extra = method;
  if (method == DH_PAIR_LEARNED || method == DH_PAIR_LEARNED_PENALTY ||
      method == DH_PAIR_PENALTY_CERT)
    return tri_pairs_learned(method, soa, n, P, Q, d, flags, st);
  if (method == DH_PAIR_MIXED)
    return tri_pairs_mixed(soa, soa32, n, eps_total, P, Q, d, flags, st);
  return set_error(DH_ERR_UNSUPPORTED, "dh_tri_pairs: method %d not available", method);
}
}  // namespace dh

extern "C" int dh_tri_pairs_counters(dh_pair_counters* out) {
  if (!out) return ::dh::set_error(DH_ERR_ARG, "null");
  out->screened = out->certified = out->features_fp64 = out->rechecked = 0;
  if (::dh::g_last_extra == DH_PAIR_MIXED) {
    unsigned long long c[2];
    int rc = ::dh::mixed_counters(c);
    if (rc) return rc;
    out->screened = static_cast<int64_t>(c[0]);
    out->rechecked = static_cast<int64_t>(c[1]);
    return DH_OK;
  }
  unsigned long long c[3];
  int rc = ::dh::learned_counters(c);
  if (rc) return rc;
  out->certified = static_cast<int64_t>(c[0]);
  out->features_fp64 = static_cast<int64_t>(c[1]);
  out->rechecked = static_cast<int64_t>(c[2]);
  return DH_OK;
}
