// ispc.hip — the reference's ISPC kernels ispc_bf (bf.ispc:3-937) and
// ispc_pen (penalty.ispc:154-174) on the GPU, behind C entry points with
// their argument shapes: n pairs in vertex-major SoA arrays (pp1..pp6[k][i]
// is coordinate k of the pair's vertex), outputs P[k][i], Q[k][i],
// distance[i].  One lane per pair (the ISPC foreach over i); the per-pair
// arithmetic is ispc_math.hpp.
#include "ispc_math.hpp"
#include "runtime.hpp"

namespace dh {

struct IspcArrays {
  const double* pp[6][3];
  double* P[3];
  double* Q[3];
  double* distance;
};

__device__ __forceinline__ void ispc_load(const IspcArrays& a, int64_t i, V3<double> p[6]) {
#pragma unroll
  for (int v = 0; v < 6; ++v) p[v] = {a.pp[v][0][i], a.pp[v][1][i], a.pp[v][2][i]};
}

// P, Q and distance are read first: the reference leaves them as they were on
// its early returns and keeps the pre-pass code in distance when the
// edge-edge minimum is not beaten.
__global__ __launch_bounds__(256) void k_ispc_bf(unsigned n, IspcArrays a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  V3<double> p[6];
  ispc_load(a, i, p);
  V3<double> P = {a.P[0][i], a.P[1][i], a.P[2][i]};
  V3<double> Q = {a.Q[0][i], a.Q[1][i], a.Q[2][i]};
  double dist = a.distance[i];
  ispc_bf(p, P, Q, dist);
  a.P[0][i] = P.x;
  a.P[1][i] = P.y;
  a.P[2][i] = P.z;
  a.Q[0][i] = Q.x;
  a.Q[1][i] = Q.y;
  a.Q[2][i] = Q.z;
  a.distance[i] = dist;
}

__global__ __launch_bounds__(256) void k_ispc_pen(unsigned n, IspcArrays a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  V3<double> p[6];
  ispc_load(a, i, p);
  const Tri<double> A = {{p[0], p[1], p[2]}}, B = {{p[3], p[4], p[5]}};
  V3<double> P, Q;
  a.distance[i] = ispc_penalty(A, B, P, Q);
  a.P[0][i] = P.x;
  a.P[1][i] = P.y;
  a.P[2][i] = P.z;
  a.Q[0][i] = Q.x;
  a.Q[1][i] = Q.y;
  a.Q[2][i] = Q.z;
}

static int ispc_args(const char* name, unsigned n, double* const* pp[6], double* const P[3],
                     double* const Q[3], double* distance, IspcArrays& a) {
  if (n == 0) return DH_OK;
  if (!P || !Q || !distance) return set_error(DH_ERR_ARG, "%s: null output", name);
  for (int v = 0; v < 6; ++v) {
    if (!pp[v]) return set_error(DH_ERR_ARG, "%s: null vertex table", name);
    for (int k = 0; k < 3; ++k) {
      if (!pp[v][k]) return set_error(DH_ERR_ARG, "%s: null coordinate array", name);
      a.pp[v][k] = pp[v][k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    if (!P[k] || !Q[k]) return set_error(DH_ERR_ARG, "%s: null output array", name);
    a.P[k] = P[k];
    a.Q[k] = Q[k];
  }
  a.distance = distance;
  return DH_OK;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_ispc_bf(unsigned n, double* const pp1[3], double* const pp2[3], double* const pp3[3],
               double* const pp4[3], double* const pp5[3], double* const pp6[3],
               double* const P[3], double* const Q[3], double* distance, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  double* const* pp[6] = {pp1, pp2, pp3, pp4, pp5, pp6};
  IspcArrays a{};
  if ((rc = ispc_args("dh_ispc_bf", n, pp, P, Q, distance, a))) return rc;
  if (n == 0) return DH_OK;
  hipLaunchKernelGGL(k_ispc_bf, grid_for(n, 256), 256, 0, as_stream(stream), n, a);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_ispc_pen(unsigned n, double* const a1[3], double* const b1[3], double* const c1[3],
                double* const d1[3], double* const e1[3], double* const f1[3],
                double* const p[3], double* const q[3], double* distance, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  double* const* pp[6] = {a1, b1, c1, d1, e1, f1};
  IspcArrays a{};
  if ((rc = ispc_args("dh_ispc_pen", n, pp, p, q, distance, a))) return rc;
  if (n == 0) return DH_OK;
  hipLaunchKernelGGL(k_ispc_pen, grid_for(n, 256), 256, 0, as_stream(stream), n, a);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

}  // extern "C"
