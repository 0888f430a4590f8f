// tri_pairs.hip — unique triangle-pair kernels (the isolated narrow phase):
// one lane per pair, 18 coalesced SoA loads, exact binary64 arithmetic of
// exact_math.hpp, and the seeded workload generator for that benchmark.
#include <cmath>

#include "exact_math.hpp"
#include "runtime.hpp"
#include "tri_pairs.hpp"
#include "zoltan_math.hpp"

namespace dh {

__device__ __forceinline__ Tri<double> load_tri(const double* __restrict__ soa, int64_t n,
                                                int64_t i, int base) {
  Tri<double> t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.v[k].x = __builtin_nontemporal_load(soa + (base + k) * n + i);
    t.v[k].y = __builtin_nontemporal_load(soa + (base + 3 + k) * n + i);
    t.v[k].z = __builtin_nontemporal_load(soa + (base + 6 + k) * n + i);
  }
  return t;
}

__device__ __forceinline__ void store_pair(int64_t n, int64_t i, const V3<double>& p,
                                           const V3<double>& q, double* __restrict__ P,
                                           double* __restrict__ Q, double* __restrict__ d) {
  __builtin_nontemporal_store(p.x, P + i);
  __builtin_nontemporal_store(p.y, P + n + i);
  __builtin_nontemporal_store(p.z, P + 2 * n + i);
  __builtin_nontemporal_store(q.x, Q + i);
  __builtin_nontemporal_store(q.y, Q + n + i);
  __builtin_nontemporal_store(q.z, Q + 2 * n + i);
  __builtin_nontemporal_store(pq_dist(p, q), d + i);
}

template <int METHOD>
__global__ __launch_bounds__(256) void k_tri_pairs_exact(const double* __restrict__ soa,
                                                         int64_t n, double max_err,
                                                         double* __restrict__ P,
                                                         double* __restrict__ Q,
                                                         double* __restrict__ d,
                                                         uint8_t* __restrict__ flags) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Tri<double> A = load_tri(soa, n, i, 0);
  const Tri<double> B = load_tri(soa, n, i, 9);
  V3<double> p, q;
  uint8_t fl = 0;
  if (METHOD == DH_PAIR_BF64) {
    if (bf_solver(A, B, p, q)) fl |= DH_FLAG_EDGE;
  } else if (METHOD == DH_PAIR_PENALTY64 || METHOD == DH_PAIR_HYBRID64) {
    const double x0[4] = {0.33, 0.33, 0.33, 0.33};
    const bool failed = penalty_fixed(A, B, x0, 4, max_err, p, q, (double*)nullptr);
    if (failed) fl |= DH_FLAG_FAILED;
    if (METHOD == DH_PAIR_HYBRID64 && failed) {
      fl |= DH_FLAG_RECHECK;
      if (bf_solver(A, B, p, q)) fl |= DH_FLAG_EDGE;
    }
  } else if (METHOD == DH_PAIR_BF64_EXACT) {
    V3<double> x = {0, 0, 0};
    const bool hit = zh_segt(B.v[0], B.v[1], A.v[0], A.v[1], A.v[2], x) == 1 ||
                     zh_segt(B.v[1], B.v[2], A.v[0], A.v[1], A.v[2], x) == 1 ||
                     zh_segt(B.v[2], B.v[0], A.v[0], A.v[1], A.v[2], x) == 1 ||
                     zh_segt(A.v[0], A.v[1], B.v[0], B.v[1], B.v[2], x) == 1 ||
                     zh_segt(A.v[1], A.v[2], B.v[0], B.v[1], B.v[2], x) == 1 ||
                     zh_segt(A.v[2], A.v[0], B.v[0], B.v[1], B.v[2], x) == 1;
    if (hit) {
      p = q = x;
      fl |= DH_FLAG_INTERSECT;
    } else if (bf_solver(A, B, p, q)) {
      fl |= DH_FLAG_EDGE;
    }
  } else {
    penalty_adaptive(A, B, max_err, p, q);
  }
  store_pair(n, i, p, q, P, Q, d);
  if (flags) flags[i] = fl;
}

// ----------------------------------------------------------------- generator
// splitmix64 (Steele, Lea, Flood 2014): counter-based uniform stream per pair.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t key;
  uint64_t ctr;
  __device__ double uniform() {  // [0, 1)
    const uint64_t v = splitmix64(key + (ctr++) * 0x9E3779B97F4A7C15ull);
    return static_cast<double>(v >> 11) * 0x1.0p-53;
  }
};

__device__ __forceinline__ V3<double> cross(const V3<double>& a, const V3<double>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3<double> axpy(double s, const V3<double>& a, const V3<double>& b) {
  return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z};
}
__device__ __forceinline__ V3<double> scale(double s, const V3<double>& a) {
  return {s * a.x, s * a.y, s * a.z};
}
__device__ __forceinline__ V3<double> normalize(const V3<double>& a) {
  return scale(1.0 / sqrt(dot(a, a)), a);
}

// does segment pq cross triangle abc (strictly)?
__device__ bool seg_crosses_tri(const V3<double>& p, const V3<double>& q, const V3<double>& a,
                                const V3<double>& b, const V3<double>& c) {
  const V3<double> n = cross(sub(b, a), sub(c, a));
  const double o1 = dot(n, sub(p, a)), o2 = dot(n, sub(q, a));
  if ((o1 > 0 && o2 > 0) || (o1 < 0 && o2 < 0)) return false;
  const V3<double> r = sub(q, p);
  const double v0 = dot(r, cross(sub(a, p), sub(b, p)));
  const double v1 = dot(r, cross(sub(b, p), sub(c, p)));
  const double v2 = dot(r, cross(sub(c, p), sub(a, p)));
  return (v0 >= 0 && v1 >= 0 && v2 >= 0) || (v0 <= 0 && v1 <= 0 && v2 <= 0);
}

__device__ bool tris_intersect(const Tri<double>& A, const Tri<double>& B) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (seg_crosses_tri(A.v[k], A.v[(k + 1) % 3], B.v[0], B.v[1], B.v[2])) return true;
    if (seg_crosses_tri(B.v[k], B.v[(k + 1) % 3], A.v[0], A.v[1], A.v[2])) return true;
  }
  return false;
}

__device__ double bf_dist(const Tri<double>& A, const Tri<double>& B) {
  V3<double> p, q;
  bf_solver(A, B, p, q);
  return pq_dist(p, q);
}

__global__ __launch_bounds__(256) void k_gen_tri_pairs(uint64_t seed, int64_t n,
                                                       const GenParams gp,
                                                       double* __restrict__ soa,
                                                       double* __restrict__ target,
                                                       uint8_t* __restrict__ kind) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Stream ra{splitmix64(seed * 2 + 1) ^ splitmix64(static_cast<uint64_t>(i)), 0};
  Stream rb{splitmix64(seed * 2 + 2) ^ splitmix64(static_cast<uint64_t>(i) + 0x1234567ull), 0};
  const double L = gp.box;
  Tri<double> A, B;
  for (int k = 0; k < 3; ++k) A.v[k] = {L * ra.uniform(), L * ra.uniform(), L * ra.uniform()};
  const double u_kind = rb.uniform();
  const double lmin = log(gp.sep_min), lmax = log(gp.sep_max);
  const double s = exp(lmin + rb.uniform() * (lmax - lmin));
  uint8_t kd;
  if (u_kind < gp.frac_parallel) {
    kd = 1;
    // an edge of B nearly parallel (angle < max_angle) to edge a0a1 of A, at
    // distance s in the edge's outward Voronoi wedge; B lies beyond it.
    const V3<double> e = sub(A.v[1], A.v[0]);
    const double len = sqrt(dot(e, e));
    const V3<double> eh = scale(1.0 / len, e);
    const V3<double> g = sub(A.v[0], A.v[2]);
    const V3<double> m = normalize(axpy(-dot(g, eh), eh, g));
    const V3<double> nA = normalize(cross(e, sub(A.v[2], A.v[0])));
    const double phi = (2.0 * rb.uniform() - 1.0) * 1.0;
    const V3<double> w = axpy(cos(phi), m, scale(sin(phi), nA));
    const double tau0 = 0.1 + 0.3 * rb.uniform();
    const V3<double> b0 = axpy(s, w, axpy(tau0, e, A.v[0]));
    const double th = gp.max_angle * rb.uniform();
    const V3<double> dir = axpy(cos(th), eh, scale(sin(th), cross(w, eh)));
    const V3<double> b1 = axpy((0.3 + 0.3 * rb.uniform()) * len, dir, b0);
    const V3<double> mid = scale(0.5, V3<double>{b0.x + b1.x, b0.y + b1.y, b0.z + b1.z});
    const V3<double> b2 = axpy((0.3 + 0.7 * rb.uniform()) * len, w, mid);
    B.v[0] = b0;
    B.v[1] = b1;
    B.v[2] = b2;
  } else {
    for (int k = 0; k < 3; ++k) B.v[k] = {L * rb.uniform(), L * rb.uniform(), L * rb.uniform()};
    // centre B on A's centroid
    V3<double> ca = {(A.v[0].x + A.v[1].x + A.v[2].x) / 3, (A.v[0].y + A.v[1].y + A.v[2].y) / 3,
                     (A.v[0].z + A.v[1].z + A.v[2].z) / 3};
    V3<double> cb = {(B.v[0].x + B.v[1].x + B.v[2].x) / 3, (B.v[0].y + B.v[1].y + B.v[2].y) / 3,
                     (B.v[0].z + B.v[1].z + B.v[2].z) / 3};
    const V3<double> sh = sub(ca, cb);
    for (int k = 0; k < 3; ++k) B.v[k] = axpy(1.0, sh, B.v[k]);
    if (u_kind < gp.frac_parallel + gp.frac_intersect) {
      kd = 2;  // pushed through A: centroids coincide
    } else {
      kd = 0;
      // uniform axis; the true distance along t*axis is convex in t and zero
      // at t = 0, so bisection on "separated and bf distance >= s" converges
      const double z = 2.0 * rb.uniform() - 1.0;
      const double ph = 6.283185307179586 * rb.uniform();
      const double rr = sqrt(fmax(0.0, 1.0 - z * z));
      const V3<double> ax = {rr * cos(ph), rr * sin(ph), z};
      double lo = 0.0, hi = s + 4.0 * L;
      for (int it = 0; it < gp.bisection_steps; ++it) {
        const double mid = 0.5 * (lo + hi);
        Tri<double> Bm;
        for (int k = 0; k < 3; ++k) Bm.v[k] = axpy(mid, ax, B.v[k]);
        const bool ok = !tris_intersect(A, Bm) && bf_dist(A, Bm) >= s;
        if (ok) hi = mid; else lo = mid;
      }
      for (int k = 0; k < 3; ++k) B.v[k] = axpy(hi, ax, B.v[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    soa[(0 + k) * n + i] = A.v[k].x;
    soa[(3 + k) * n + i] = A.v[k].y;
    soa[(6 + k) * n + i] = A.v[k].z;
    soa[(9 + k) * n + i] = B.v[k].x;
    soa[(12 + k) * n + i] = B.v[k].y;
    soa[(15 + k) * n + i] = B.v[k].z;
  }
  if (target) target[i] = s;
  if (kind) kind[i] = kd;
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = static_cast<float>(a[i]);
}

}  // namespace dh

using namespace dh;

namespace dh {
int tri_pairs_extra(int method, const double* soa, const float* soa32, int64_t n,
                    double eps_total, double* P, double* Q, double* d, uint8_t* flags,
                    hipStream_t st);
}

extern "C" {

int dh_tri_pairs(int method, const double* soa, const float* soa32, int64_t n, double eps_total,
                 double* P, double* Q, double* d, uint8_t* flags, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!soa || !P || !Q || !d)))
    return set_error(DH_ERR_ARG, "dh_tri_pairs: bad arguments");
  if (n == 0) return DH_OK;
  const hipStream_t st = as_stream(stream);
  const double max_err = eps_total / 16.0;
  const int block = tuning().exact_block;
  const unsigned grid = grid_for(n, block);
  switch (method) {
    case DH_PAIR_BF64:
      hipLaunchKernelGGL(k_tri_pairs_exact<DH_PAIR_BF64>, grid, block, 0, st, soa, n, max_err, P,
                         Q, d, flags);
      break;
    case DH_PAIR_PENALTY64:
      hipLaunchKernelGGL(k_tri_pairs_exact<DH_PAIR_PENALTY64>, grid, block, 0, st, soa, n,
                         max_err, P, Q, d, flags);
      break;
    case DH_PAIR_HYBRID64:
      hipLaunchKernelGGL(k_tri_pairs_exact<DH_PAIR_HYBRID64>, grid, block, 0, st, soa, n, max_err,
                         P, Q, d, flags);
      break;
    case DH_PAIR_PENALTY_ADAPT64:
      hipLaunchKernelGGL(k_tri_pairs_exact<DH_PAIR_PENALTY_ADAPT64>, grid, block, 0, st, soa, n,
                         max_err, P, Q, d, flags);
      break;
    case DH_PAIR_BF64_EXACT:
      hipLaunchKernelGGL(k_tri_pairs_exact<DH_PAIR_BF64_EXACT>, grid, block, 0, st, soa, n, max_err,
                         P, Q, d, flags);
      break;
    case DH_PAIR_MIXED:
    case DH_PAIR_LEARNED:
    case DH_PAIR_LEARNED_PENALTY:
    case DH_PAIR_PENALTY_CERT:
      return tri_pairs_extra(method, soa, soa32, n, eps_total, P, Q, d, flags, st);
    default:
      return set_error(DH_ERR_ARG, "dh_tri_pairs: unknown method %d", method);
  }
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_gen_tri_pairs(uint64_t seed, int64_t n, const dh_gen_params* p, double* soa_dev,
                     double* target_dev, uint8_t* kind_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || (n > 0 && !soa_dev)) return set_error(DH_ERR_ARG, "dh_gen_tri_pairs: bad args");
  if (n == 0) return DH_OK;
  GenParams gp;
  gp.box = p ? p->box : 1e-3;
  gp.sep_min = p ? p->sep_min : 1e-7;
  gp.sep_max = p ? p->sep_max : 1e-2;
  gp.frac_parallel = p ? p->frac_parallel : 0.05;
  gp.frac_intersect = p ? p->frac_intersect : 0.01;
  gp.max_angle = p ? p->max_angle : 1e-4;
  gp.bisection_steps = p ? p->bisection_steps : 48;
  hipLaunchKernelGGL(k_gen_tri_pairs, grid_for(n, 256), 256, 0, as_stream(stream), seed, n, gp,
                     soa_dev, target_dev, kind_dev);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_convert_f32(const double* src_dev, float* dst_dev, int64_t n, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n <= 0) return DH_OK;
  hipLaunchKernelGGL(k_to_f32, grid_for(n, 256), 256, 0, as_stream(stream), src_dev, dst_dev, n);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

}  // extern "C"
