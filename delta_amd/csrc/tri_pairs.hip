// tri_pairs.hip — unique triangle-pair kernels (the isolated narrow phase):
// one lane per pair, 18 coalesced SoA loads, exact binary64 arithmetic of
// exact_math.hpp, and the seeded workload generator for that benchmark.
#include <cmath>

#include "exact_math.hpp"
#include "gen_pairs.hpp"
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
// One lane per pair (gen_pairs.hpp); pair i writes only its own column i of
// the 18 x n SoA block and entry i of target and kind.
__global__ __launch_bounds__(256) void k_gen_tri_pairs(uint64_t seed, int64_t n,
                                                       const GenParams gp,
                                                       double* __restrict__ soa,
                                                       double* __restrict__ target,
                                                       uint8_t* __restrict__ kind) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Tri<double> A, B;
  double s;
  uint8_t kd;
  gen_pair(seed, i, gp, A, B, s, kd);
  for (int k = 0; k < 3; ++k) {
    soa[soa_index(0 + k, n, i)] = A.v[k].x;
    soa[soa_index(3 + k, n, i)] = A.v[k].y;
    soa[soa_index(6 + k, n, i)] = A.v[k].z;
    soa[soa_index(9 + k, n, i)] = B.v[k].x;
    soa[soa_index(12 + k, n, i)] = B.v[k].y;
    soa[soa_index(15 + k, n, i)] = B.v[k].z;
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
