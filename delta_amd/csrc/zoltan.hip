// zoltan.hip — the reference's zoltan-application contact kernels on the GPU
// (zoltan/{bf,penalty,hybrid,contact}.cpp), behind C entry points with the
// reference's argument shapes (vertex-major SoA arrays d[k][i]).
//
//   dh_zoltan_bf         one lane per pair, zoltan_bf (zoltan_math.hpp)
//   dh_zoltan_penalty    one lane per pair i in [it, nt)
//   dh_zoltan_hybrid     one lane per batch of 4 pairs, batches [it, nt/4): the
//                        reference's sequential semantics (its OpenMP loops race
//                        on shared A..F): 4 penalties, mean error > 1e-8 ->
//                        zh_bf on the batch's last pair, written to all four
//   dh_zoltan_detection  all-to-all or two-range triangle detection of
//                        zoltan/contact.cpp, in three passes: per-row candidate
//                        counts, ordered candidate rows, then the per-particle
//                        first-come de-duplication and the ordered compaction
#include "runtime.hpp"
#include "zoltan_math.hpp"

#include <vector>

namespace dh {

struct ZArrays {
  const double* d[3];
  const double* e[3];
  const double* f[3];
  double* p[3];
  double* q[3];
};

__device__ __forceinline__ V3<double> zload(const double* const a[3], int64_t i) {
  return {a[0][i], a[1][i], a[2][i]};
}

__device__ __forceinline__ void zstore(double* const a[3], int64_t i, const V3<double>& v) {
  a[0][i] = v.x;
  a[1][i] = v.y;
  a[2][i] = v.z;
}

__global__ __launch_bounds__(256) void k_zoltan_bf(const double* __restrict__ soa, int64_t n,
                                                   double* __restrict__ P, double* __restrict__ Q,
                                                   double* __restrict__ d,
                                                   uint8_t* __restrict__ flags) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  Tri<double> A, B;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    A.v[k] = {soa[k * n + i], soa[(3 + k) * n + i], soa[(6 + k) * n + i]};
    B.v[k] = {soa[(9 + k) * n + i], soa[(12 + k) * n + i], soa[(15 + k) * n + i]};
  }
  V3<double> p, q;
  const int r = zoltan_bf(A, B, p, q);
  if (r >= 0) {
    P[i] = p.x;
    P[n + i] = p.y;
    P[2 * n + i] = p.z;
    Q[i] = q.x;
    Q[n + i] = q.y;
    Q[2 * n + i] = q.z;
    if (d) d[i] = pq_dist(p, q);
  }
  if (flags) flags[i] = r < 0 ? DH_FLAG_FAILED : (r == 1 ? DH_FLAG_EDGE : 0);
}

__global__ __launch_bounds__(256) void k_zoltan_penalty(unsigned it, unsigned nt, Tri<double> A,
                                                        ZArrays z) {
  const unsigned i = it + blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  Tri<double> B;
  B.v[0] = zload(z.d, i);
  B.v[1] = zload(z.e, i);
  B.v[2] = zload(z.f, i);
  V3<double> P, Q;
  zoltan_penalty(A, B, i, P, Q);
  zstore(z.p, i, P);
  zstore(z.q, i, Q);
}

__global__ __launch_bounds__(256) void k_zoltan_hybrid(unsigned k0, unsigned k1, Tri<double> A,
                                                       ZArrays z) {
  const unsigned k = k0 + blockIdx.x * 256 + threadIdx.x;
  if (k >= k1) return;
  double err[4];
  Tri<double> B;
  V3<double> P, Q;
  for (unsigned l = 0; l < 4; ++l) {
    const unsigned i = k * 4 + l;
    B.v[0] = zload(z.d, i);
    B.v[1] = zload(z.e, i);
    B.v[2] = zload(z.f, i);
    err[l] = zh_penalty(A, B, P, Q);
    zstore(z.p, i, P);
    zstore(z.q, i, Q);
  }
  double batch = 0.0;
  for (int l = 0; l < 4; ++l) batch += err[l];
  if (batch / 4u > 1E-8) {
    for (unsigned l = 0; l < 4; ++l) {
      zh_bf(A.v[0], A.v[1], A.v[2], B.v[0], B.v[1], B.v[2], P, Q);
      zstore(z.p, k * 4 + l, P);
      zstore(z.q, k * 4 + l, Q);
    }
  }
}

// ------------------------------------------------------ zoltan detection
constexpr double kZoltanMargin = 15e-2;  // zoltan/contact.cpp:95

struct ZDetArgs {
  const double* t[3][3];  // t[vertex][axis][triangle]
  const int* pid;
  int s1, e1, s2, e2;
  bool all;  // all-to-all over [s1, e1): j > i
  const long long* row_off;
  int* row_count;
  dh_zoltan_contact* cand;
};

__device__ __forceinline__ Tri<double> ztri(const ZDetArgs& a, int i) {
  Tri<double> r;
#pragma unroll
  for (int v = 0; v < 3; ++v) r.v[v] = {a.t[v][0][i], a.t[v][1][i], a.t[v][2][i]};
  return r;
}

// The contact of pair (i, j) if its distance is below the margin.
__device__ __forceinline__ bool zpair(const ZDetArgs& a, int i, int j, dh_zoltan_contact* c) {
  if (a.pid[i] == a.pid[j]) return false;
  V3<double> P, Q;
  if (zoltan_bf(ztri(a, i), ztri(a, j), P, Q) < 0) return false;
  const double dx = Q.x - P.x, dy = Q.y - P.y, dz = Q.z - P.z;
  const double dist = dsqrt(dx * dx + dy * dy + dz * dz);
  if (!(dist < kZoltanMargin)) return false;
  if (c) {
    const double depth = kZoltanMargin - dist;
    c->pid[0] = a.pid[i];
    c->pid[1] = a.pid[j];
    c->color[0] = c->color[1] = 0;
    c->point[0] = (P.x + Q.x) / 2;
    c->point[1] = (P.y + Q.y) / 2;
    c->point[2] = (P.z + Q.z) / 2;
    c->depth = depth;
    c->normal[0] = (Q.x - P.x) / depth;
    c->normal[1] = (Q.y - P.y) / depth;
    c->normal[2] = (Q.z - P.z) / depth;
    c->pp[0] = P.x;
    c->pp[1] = P.y;
    c->pp[2] = P.z;
    c->qq[0] = Q.x;
    c->qq[1] = Q.y;
    c->qq[2] = Q.z;
  }
  return true;
}

// One workgroup per row i; WRITE = false counts, true writes the row's
// candidates in j order at row_off[i - s1].
template <bool WRITE>
__global__ __launch_bounds__(256) void k_zdet_rows(ZDetArgs a) {
  __shared__ int wave_n[4];
  __shared__ int running;
  const int i = a.s1 + blockIdx.x;
  const int jb = a.all ? i + 1 : a.s2, je = a.all ? a.e1 : a.e2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) running = 0;
  __syncthreads();
  for (int j0 = jb; j0 < je; j0 += 256) {
    const int j = j0 + threadIdx.x;
    dh_zoltan_contact c;
    const bool hit = j < je && zpair(a, i, j, WRITE ? &c : nullptr);
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wave_n[wave] = __popcll(m);
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wave_n[w];
    const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(m >> 32),
                                              __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0));
    if (WRITE && hit) a.cand[a.row_off[blockIdx.x] + running + before + pre] = c;
    __syncthreads();
    if (threadIdx.x == 0) running += wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
    __syncthreads();
  }
  if (!WRITE && threadIdx.x == 0) a.row_count[blockIdx.x] = running;
}

__device__ __forceinline__ bool zsame_any(const dh_zoltan_contact& a, const dh_zoltan_contact& b) {
  return a.pp[0] == b.pp[0] || a.pp[1] == b.pp[1] || a.pp[2] == b.pp[2] || a.qq[0] == b.qq[0] ||
         a.qq[1] == b.qq[1] || a.qq[2] == b.qq[2];
}

// zoltan/contact.cpp:100-108: a candidate owned by particle p is dropped when
// any coordinate of its P or Q equals that of a contact p already holds.  One
// lane per particle walks the candidates in discovery order.
__global__ __launch_bounds__(256) void k_zdet_dedup(const dh_zoltan_contact* __restrict__ cand,
                                                    long long nc, int npid,
                                                    int* __restrict__ keep) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= npid) return;
  for (long long c = 0; c < nc; ++c) {
    if (cand[c].pid[0] != p) continue;
    bool found = false;
    for (long long o = 0; o < c && !found; ++o)
      if (keep[o] && cand[o].pid[0] == p && zsame_any(cand[o], cand[c])) found = true;
    keep[c] = found ? 0 : 1;
  }
}

__global__ __launch_bounds__(256) void k_zdet_compact(const dh_zoltan_contact* __restrict__ cand,
                                                      const int* __restrict__ keep,
                                                      const long long* __restrict__ offs,
                                                      long long nc, dh_zoltan_contact* out,
                                                      long long cap) {
  const long long c = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (c >= nc || !keep[c]) return;
  const long long o = offs[c];
  if (o < cap) out[o] = cand[c];
}

int scan_counts(const int* counts, int n, long long* offs, hipStream_t st);

struct ZoltanScratch {
  DevBuf row_count, row_off, cand, keep, keep_off;
};
static ZoltanScratch g_z;

static int zoltan_detection(int s1, int e1, int s2, int e2, bool all, double* const t[6][3],
                            const int* pid, int npid, dh_zoltan_contact* out, int cap, int* n_out,
                            hipStream_t st) {
  *n_out = 0;
  const int rows = e1 - s1;
  if (rows <= 0) return DH_OK;
  ZDetArgs a;
  for (int v = 0; v < 3; ++v)
    for (int x = 0; x < 3; ++x) a.t[v][x] = t[v][x];
  a.pid = pid;
  a.s1 = s1;
  a.e1 = e1;
  a.s2 = s2;
  a.e2 = e2;
  a.all = all;
  int rc;
  if ((rc = g_z.row_count.grow(sizeof(int) * (rows + 1)))) return rc;
  if ((rc = g_z.row_off.grow(sizeof(long long) * (rows + 1)))) return rc;
  a.row_count = g_z.row_count.as<int>();
  a.row_off = g_z.row_off.as<long long>();
  a.cand = nullptr;
  hipLaunchKernelGGL(k_zdet_rows<false>, rows, 256, 0, st, a);
  if ((rc = scan_counts(a.row_count, rows, g_z.row_off.as<long long>(), st))) return rc;
  long long nc = 0;
  DH_HIP_CHECK(hipMemcpyAsync(&nc, g_z.row_off.as<long long>() + rows, sizeof(nc),
                              hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  if (nc == 0) return DH_OK;
  if (nc > 0x7fffffffLL) return set_error(DH_ERR_CAPACITY, "zoltan detection: too many candidates");
  if ((rc = g_z.cand.grow(sizeof(dh_zoltan_contact) * nc))) return rc;
  if ((rc = g_z.keep.grow(sizeof(int) * (nc + 1)))) return rc;
  if ((rc = g_z.keep_off.grow(sizeof(long long) * (nc + 1)))) return rc;
  a.cand = g_z.cand.as<dh_zoltan_contact>();
  // candidates whose owner is outside [0, npid) stay dropped
  DH_HIP_CHECK(hipMemsetAsync(g_z.keep.ptr, 0, sizeof(int) * (nc + 1), st));
  hipLaunchKernelGGL(k_zdet_rows<true>, rows, 256, 0, st, a);
  hipLaunchKernelGGL(k_zdet_dedup, grid_for(npid, 256), 256, 0, st, a.cand, nc, npid,
                     g_z.keep.as<int>());
  if ((rc = scan_counts(g_z.keep.as<int>(), static_cast<int>(nc), g_z.keep_off.as<long long>(), st)))
    return rc;
  long long kept = 0;
  DH_HIP_CHECK(hipMemcpyAsync(&kept, g_z.keep_off.as<long long>() + nc, sizeof(kept),
                              hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  *n_out = static_cast<int>(kept);
  if (kept > cap) return DH_ERR_CAPACITY;
  hipLaunchKernelGGL(k_zdet_compact, grid_for(nc, 256), 256, 0, st, a.cand, g_z.keep.as<int>(),
                     g_z.keep_off.as<long long>(), nc, out, static_cast<long long>(cap));
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

static int zarrays(double* const d[3], double* const e[3], double* const f[3], double* const p[3],
                   double* const q[3], ZArrays& z) {
  if (!d || !e || !f || !p || !q) return set_error(DH_ERR_ARG, "zoltan: null array table");
  for (int k = 0; k < 3; ++k) {
    if (!d[k] || !e[k] || !f[k] || !p[k] || !q[k]) return set_error(DH_ERR_ARG, "zoltan: null array");
    z.d[k] = d[k];
    z.e[k] = e[k];
    z.f[k] = f[k];
    z.p[k] = p[k];
    z.q[k] = q[k];
  }
  return DH_OK;
}

static Tri<double> tri_of(const double a[3], const double b[3], const double c[3]) {
  Tri<double> t;
  t.v[0] = {a[0], a[1], a[2]};
  t.v[1] = {b[0], b[1], b[2]};
  t.v[2] = {c[0], c[1], c[2]};
  return t;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_zoltan_bf(const double* soa_dev, int64_t n, double* P_dev, double* Q_dev, double* d_dev,
                 uint8_t* flags_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!soa_dev || !P_dev || !Q_dev)))
    return set_error(DH_ERR_ARG, "dh_zoltan_bf: bad arguments");
  if (n > 0)
    hipLaunchKernelGGL(k_zoltan_bf, grid_for(n, 256), 256, 0, as_stream(stream), soa_dev, n, P_dev,
                       Q_dev, d_dev, flags_dev);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_zoltan_penalty(unsigned it, unsigned nt, const double a[3], const double b[3],
                      const double c[3], double* const d[3], double* const e[3],
                      double* const f[3], double* const p[3], double* const q[3], void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!a || !b || !c) return set_error(DH_ERR_ARG, "dh_zoltan_penalty: null triangle");
  ZArrays z;
  if ((rc = zarrays(d, e, f, p, q, z))) return rc;
  if (nt > it)
    hipLaunchKernelGGL(k_zoltan_penalty, grid_for(nt - it, 256), 256, 0, as_stream(stream), it, nt,
                       tri_of(a, b, c), z);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_zoltan_hybrid(int it, int nt, const double a[3], const double b[3], const double c[3],
                     double* const d[3], double* const e[3], double* const f[3],
                     double* const p[3], double* const q[3], void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!a || !b || !c || it < 0 || nt < 0)
    return set_error(DH_ERR_ARG, "dh_zoltan_hybrid: bad arguments");
  ZArrays z;
  if ((rc = zarrays(d, e, f, p, q, z))) return rc;
  const unsigned k0 = static_cast<unsigned>(it), k1 = static_cast<unsigned>(nt) / 4u;
  if (k1 > k0)
    hipLaunchKernelGGL(k_zoltan_hybrid, grid_for(k1 - k0, 256), 256, 0, as_stream(stream), k0, k1,
                       tri_of(a, b, c), z);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int dh_zoltan_detection(int s, int e, double* const t[6][3], const int* tid_dev,
                        const int* pid_dev, int npid, dh_zoltan_contact* out_dev, int cap,
                        int* n_out, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  (void)tid_dev;
  if (!t || !pid_dev || !n_out || npid < 0 || cap < 0 || s < 0 || e < s)
    return set_error(DH_ERR_ARG, "dh_zoltan_detection: bad arguments");
  return zoltan_detection(s, e, 0, 0, true, t, pid_dev, npid, out_dev, cap, n_out,
                          as_stream(stream));
}

int dh_zoltan_detection2(int s1, int e1, int s2, int e2, double* const t[6][3],
                         const int* tid_dev, const int* pid_dev, int npid,
                         dh_zoltan_contact* out_dev, int cap, int* n_out, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  (void)tid_dev;
  if (!t || !pid_dev || !n_out || npid < 0 || cap < 0 || s1 < 0 || e1 < s1 || s2 < 0 || e2 < s2)
    return set_error(DH_ERR_ARG, "dh_zoltan_detection2: bad arguments");
  return zoltan_detection(s1, e1, s2, e2, false, t, pid_dev, npid, out_dev, cap, n_out,
                          as_stream(stream));
}

}  // extern "C"
