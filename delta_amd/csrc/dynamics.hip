// dynamics.hip — contact forces and the fused explicit update.
//
// k_pair_forces:  one lane per particle pair, the pair's contacts in order,
//                 both sides (A with the stored normal, B with it negated).
// k_particle_sum: one lane per particle, its pairs in ascending partner index
//                 (deterministic, so the step is bit-reproducible).
// k_integrate:    one wave per particle; every lane recomputes the particle's
//                 velocity/rotation update (identical arithmetic), lane 0 stores
//                 it, and the wave rewrites the particle's vertex replicas.
#include "dyn_math.hpp"
#include "runtime.hpp"

namespace dh {

struct BodiesDev {
  double* pos;
  double* refpos;
  double* linvel;
  double* angvel;
  double* refangvel;
  double* rot;
  double* inertia;
  double* inverse;
  double* mass;
  int* obstacle;
  int n;
};

static BodiesDev to_dev(const dh_bodies* b) {
  return {b->pos,  b->refpos,  b->linvel, b->angvel,   b->refangvel, b->rot,
          b->inertia, b->inverse, b->mass, b->obstacle, b->n};
}

__device__ __forceinline__ void load3(const double* p, int i, double v[3]) {
  v[0] = p[3 * i];
  v[1] = p[3 * i + 1];
  v[2] = p[3 * i + 2];
}

// np_dev (may be null): the live pair count when npairs is a capacity.
__device__ __forceinline__ bool past_pairs(int p, int npairs, const long long* np_dev) {
  return p >= npairs || (np_dev && p >= *np_dev);
}

// Pair p's contact forces on both sides into pf, and its incidence counts
// when it carries contacts.
__device__ __forceinline__ void pair_forces_one(const BodiesDev& b, const int* __restrict__ pairs,
                                                int p, const dh_contact* __restrict__ cs,
                                                const long long* __restrict__ off,
                                                const ForceConsts& k, double* __restrict__ pf,
                                                int* __restrict__ incidence, long long ncap) {
  const int a = pairs[2 * p], c = pairs[2 * p + 1];
  double posA[3], angA[3], linA[3], posB[3], angB[3], linB[3];
  load3(b.pos, a, posA);
  load3(b.angvel, a, angA);
  load3(b.linvel, a, linA);
  load3(b.pos, c, posB);
  load3(b.angvel, c, angB);
  load3(b.linvel, c, linB);
  const double mA = b.mass[a], mB = b.mass[c];
  double fA[3] = {0, 0, 0}, tA[3] = {0, 0, 0}, fB[3] = {0, 0, 0}, tB[3] = {0, 0, 0};
  // contacts past the capacity ncap were dropped (reported by the caller)
  const long long c0 = off[p] < ncap ? off[p] : ncap, c1 = off[p + 1] < ncap ? off[p + 1] : ncap;
  for (long long i = c0; i < c1; ++i) {
    const dh_contact& ct = cs[i];
    const double n[3] = {ct.normal[0], ct.normal[1], ct.normal[2]};
    const double nn[3] = {-1.0 * ct.normal[0], -1.0 * ct.normal[1], -1.0 * ct.normal[2]};
    contact_force(ct.x, n, ct.depth, ct.friction != 0, posA, angA, linA, mA, posB, angB, linB, mB,
                  k, fA, tA);
    contact_force(ct.x, nn, ct.depth, ct.friction != 0, posB, angB, linB, mB, posA, angA, linA,
                  mA, k, fB, tB);
  }
  double* o = pf + 12 * static_cast<long long>(p);
  for (int q = 0; q < 3; ++q) {
    o[q] = fA[q];
    o[3 + q] = tA[q];
    o[6 + q] = fB[q];
    o[9 + q] = tB[q];
  }
  if (c1 > c0) {
    atomicAdd(&incidence[a], 1);
    atomicAdd(&incidence[c], 1);
  }
}

__global__ __launch_bounds__(256) void k_pair_forces(BodiesDev b, const int* __restrict__ pairs,
                                                     int npairs,
                                                     const dh_contact* __restrict__ cs,
                                                     const long long* __restrict__ off, ForceConsts k,
                                                     double* __restrict__ pf,
                                                     int* __restrict__ incidence,
                                                     const long long* __restrict__ np_dev,
                                                     long long ncap) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (past_pairs(p, npairs, np_dev)) return;
  pair_forces_one(b, pairs, p, cs, off, k, pf, incidence, ncap);
}

// Pair p's two incidence entries, when it carries contacts (the same test as
// pair_forces_one's count, offsets clamped to ncap).
__device__ __forceinline__ void fill_one(const int* __restrict__ pairs, int p,
                                         const long long* __restrict__ off,
                                         const long long* __restrict__ start,
                                         int* __restrict__ fill, int* __restrict__ entries,
                                         long long ncap) {
  const long long c0 = off[p] < ncap ? off[p] : ncap, c1 = off[p + 1] < ncap ? off[p + 1] : ncap;
  if (c1 <= c0) return;
  const int a = pairs[2 * p], c = pairs[2 * p + 1];
  const int ia = atomicAdd(&fill[a], 1);
  entries[start[a] + ia] = 2 * p;  // side A
  const int ic = atomicAdd(&fill[c], 1);
  entries[start[c] + ic] = 2 * p + 1;  // side B
}

// CSR of contact-carrying pairs per particle: entries (pair, side) appended in
// arbitrary order, ordered by partner in k_particle_sum.
__global__ __launch_bounds__(256) void k_fill_incidence(const int* __restrict__ pairs, int npairs,
                                                        const long long* __restrict__ off,
                                                        const long long* __restrict__ start,
                                                        int* __restrict__ fill,
                                                        int* __restrict__ entries,
                                                        const long long* __restrict__ np_dev,
                                                        long long ncap) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (past_pairs(p, npairs, np_dev)) return;
  fill_one(pairs, p, off, start, fill, entries, ncap);
}

// Particles with more than this many contact-carrying pairs (walls, floors)
// are summed by a workgroup each (k_particle_sum_wide); the rest by one lane.
constexpr int kNarrowSum = 64;
constexpr int kWideSort = 4096;  // entries sorted in LDS

__device__ __forceinline__ int entry_partner(const int* __restrict__ pairs, int en) {
  const int pr = en >> 1;
  return (en & 1) ? pairs[2 * pr] : pairs[2 * pr + 1];
}

__device__ __forceinline__ const double* entry_forces(const double* __restrict__ pf, int en) {
  return pf + 12 * static_cast<long long>(en >> 1) + ((en & 1) ? 6 : 0);
}

// Sum of particle i's entries in ascending partner index by repeated
// selection of the next larger partner (m small).
__device__ void sum_by_selection(const int* __restrict__ pairs, const int* __restrict__ entries,
                                 const double* __restrict__ pf, long long s, int m, double f[3],
                                 double t[3]) {
  int last = -1;
  for (int done = 0; done < m; ++done) {
    int best = -1, bestp = 0x7fffffff;
    for (int e = 0; e < m; ++e) {
      const int en = entries[s + e];
      const int partner = entry_partner(pairs, en);
      if (partner > last && partner < bestp) {
        bestp = partner;
        best = en;
      }
    }
    if (best < 0) break;  // no larger partner: inconsistent counts, never read outside pf
    last = bestp;
    const double* o = entry_forces(pf, best);
    for (int q = 0; q < 3; ++q) {
      f[q] += o[q];
      t[q] += o[3 + q];
    }
  }
}

__global__ __launch_bounds__(256) void k_particle_sum(int n, const int* __restrict__ pairs,
                                                      const long long* __restrict__ start,
                                                      const int* __restrict__ count,
                                                      const int* __restrict__ entries,
                                                      const double* __restrict__ pf,
                                                      double* __restrict__ force,
                                                      double* __restrict__ torque,
                                                      int* __restrict__ touched,
                                                      int* __restrict__ wide) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int m = i < n ? count[i] : 0;
  // listed for k_particle_sum_wide (wide[0] entries from wide[1]), each
  // wave's in particle order
  const bool is_wide = m > kNarrowSum;
  const unsigned long long wm = __ballot(is_wide);
  if (wm) {
    const int lane = static_cast<int>(threadIdx.x & 63), first = __builtin_ctzll(wm);
    int base = 0;
    if (lane == first) base = atomicAdd(&wide[0], __popcll(wm));
    base = __shfl(base, first);
    const int pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(wm >> 32),
                                              __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(wm), 0));
    if (is_wide) wide[1 + base + pre] = i;
  }
  if (i >= n || is_wide) return;
  double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
  sum_by_selection(pairs, entries, pf, start[i], m, f, t);
  for (int q = 0; q < 3; ++q) {
    force[3 * i + q] = f[q];
    torque[3 * i + q] = t[q];
  }
  if (touched) touched[i] = m > 0 ? 1 : 0;
}

// Particle i's sum over its m > kNarrowSum entries, by the NT threads of a
// workgroup: entries sorted by partner in LDS (a bitonic sort of keys
// partner << 32 | position, so a repeated pair's by position), the entries'
// force rows staged 256 at a time, and six lanes adding them, one component
// each, in ascending partner order: the same order and arithmetic as the
// narrow sum.  Over kWideSort entries, one lane selects them in order.
template <int NT>
__device__ void wide_sum_one(int i, const int* __restrict__ pairs,
                             const long long* __restrict__ start,
                             const int* __restrict__ count, const int* __restrict__ entries,
                             const double* __restrict__ pf, double* __restrict__ force,
                             double* __restrict__ torque, int* __restrict__ touched,
                             long long* key, double* stage) {
  const int tid = threadIdx.x;
  const int m = count[i];
  const long long s = start[i];
  if (m > kWideSort) {
    if (tid == 0) {
      double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
      sum_by_selection(pairs, entries, pf, s, m, f, t);
      for (int q = 0; q < 3; ++q) {
        force[3 * i + q] = f[q];
        torque[3 * i + q] = t[q];
      }
    }
  } else {
    int P = 1;
    while (P < m) P <<= 1;
    for (int e = tid; e < P; e += NT)
      key[e] = e < m ? (static_cast<long long>(entry_partner(pairs, entries[s + e])) << 32) | e
                     : 0x7fffffffffffffffll;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
      for (int h = k >> 1; h > 0; h >>= 1) {
        for (int e = tid; e < P; e += NT) {
          const int o = e ^ h;
          if (o > e) {
            const long long a = key[e], b = key[o];
            if ((a > b) == ((e & k) == 0)) {
              key[e] = b;
              key[o] = a;
            }
          }
        }
        __syncthreads();
      }
    // lane c < 6 adds component c (force x, y, z, torque x, y, z) of every
    // row in ascending partner order
    double acc = 0.0;
    for (int c0 = 0; c0 < m; c0 += 256) {
      const int r = c0 + tid;
      if (tid < 256 && r < m) {
        const double* o = entry_forces(pf, entries[s + static_cast<int>(key[r] & 0xffffffffll)]);
        for (int q = 0; q < 6; ++q) stage[6 * tid + q] = o[q];
      }
      __syncthreads();
      if (tid < 6) {
        const int nr = m - c0 < 256 ? m - c0 : 256;
#pragma unroll 16
        for (int k = 0; k < nr; ++k) acc += stage[6 * k + tid];
      }
      __syncthreads();
    }
    if (tid < 3) force[3 * i + tid] = acc;
    else if (tid < 6) torque[3 * i + tid - 3] = acc;
  }
  if (tid == 0 && touched) touched[i] = 1;
  __syncthreads();  // the next particle reuses the keys and the staging rows
}

// The particles with more than kNarrowSum entries (listed by k_particle_sum),
// one workgroup at a time.
__global__ __launch_bounds__(256) void k_particle_sum_wide(const int* __restrict__ wide,
                                                           const int* __restrict__ pairs,
                                                           const long long* __restrict__ start,
                                                           const int* __restrict__ count,
                                                           const int* __restrict__ entries,
                                                           const double* __restrict__ pf,
                                                           double* __restrict__ force,
                                                           double* __restrict__ torque,
                                                           int* __restrict__ touched) {
  __shared__ long long key[kWideSort];
  __shared__ double stage[256 * 6];
  const int nwide = wide[0];
  for (int j = blockIdx.x; j < nwide; j += gridDim.x)
    wide_sum_one<256>(wide[1 + j], pairs, start, count, entries, pf, force, torque, touched, key,
                      stage);
}

// Particles, pairs and contact capacity for which k_forces_small does the
// force stage in one workgroup.
constexpr int kForcesSmallN = 1024;

// The whole force stage of a small scene (at most kForcesSmallN particles) in
// one workgroup, phase by phase: the pairs' forces and incidence counts, the
// exclusive scan of the counts, the incidence entries, the narrow sums, then
// the wide particles one after another.  Each phase is the multi-launch
// path's own code, so the result is the same bit for bit.  The counts, fill
// cursors and offsets live in LDS (its atomics are the workgroup's own); the
// global rows and entries are each written once and first read after a
// barrier.
__global__ __launch_bounds__(1024) void k_forces_small(
    BodiesDev b, const int* __restrict__ pairs, int npairs, const long long* __restrict__ np_dev,
    const dh_contact* __restrict__ cs, const long long* __restrict__ off, ForceConsts k,
    long long ncap, double* __restrict__ pf, int* __restrict__ entries,
    double* __restrict__ force, double* __restrict__ torque, int* __restrict__ touched) {
  __shared__ long long key[kWideSort];  // the scan's buffer first, then the sort keys
  __shared__ double stage[256 * 6];
  __shared__ long long start[kForcesSmallN + 1];
  __shared__ int cnt[kForcesSmallN];
  __shared__ int fill[kForcesSmallN];  // fill cursors, then the wide particles
  __shared__ int nwide;
  const int n = b.n, tid = threadIdx.x;
  const int live = np_dev && *np_dev < npairs ? static_cast<int>(*np_dev) : npairs;
  if (tid < kForcesSmallN) {
    cnt[tid] = 0;
    fill[tid] = 0;
  }
  if (tid == 0) nwide = 0;
  __syncthreads();
  for (int p = tid; p < live; p += 1024) pair_forces_one(b, pairs, p, cs, off, k, pf, cnt, ncap);
  __syncthreads();
  // start[0..n) exclusive, start[n] the total
  const long long v = tid < n ? cnt[tid] : 0;
  key[tid] = v;
  __syncthreads();
  for (int sh = 1; sh < 1024; sh <<= 1) {
    const long long t = tid >= sh ? key[tid - sh] : 0;
    __syncthreads();
    key[tid] += t;
    __syncthreads();
  }
  if (tid < n) start[tid] = key[tid] - v;
  if (tid == 1023) start[n] = key[1023];
  __syncthreads();
  for (int p = tid; p < live; p += 1024) fill_one(pairs, p, off, start, fill, entries, ncap);
  __syncthreads();
  int* wide = fill;  // the cursors are spent
  for (int i = tid; i < n; i += 1024) {
    const int m = cnt[i];
    if (m > kNarrowSum) {
      wide[atomicAdd(&nwide, 1)] = i;
      continue;
    }
    double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
    sum_by_selection(pairs, entries, pf, start[i], m, f, t);
    for (int q = 0; q < 3; ++q) {
      force[3 * i + q] = f[q];
      torque[3 * i + q] = t[q];
    }
    if (touched) touched[i] = m > 0 ? 1 : 0;
  }
  __syncthreads();
  const int nw = nwide;
  for (int j = 0; j < nw; ++j)
    wide_sum_one<1024>(wide[j], pairs, start, cnt, entries, pf, force, torque, touched, key,
                       stage);
}

struct Gravity {
  double g[3];
};

constexpr int kIntegrateBatch = 8;  // vertex replicas per lane per pass of k_integrate

__global__ __launch_bounds__(64) void k_integrate(BodiesDev b, const double* __restrict__ force,
                                                  const double* __restrict__ torque,
                                                  const int* __restrict__ touched, double h_arg,
                                                  const double* __restrict__ h_dev,
                                                  Gravity grav, double* __restrict__ x,
                                                  double* __restrict__ y, double* __restrict__ z,
                                                  const double* __restrict__ rx,
                                                  const double* __restrict__ ry,
                                                  const double* __restrict__ rz,
                                                  const int* __restrict__ tri_offset) {
  const int i = blockIdx.x;
  const int lane = static_cast<int>(threadIdx.x);
  const double h = h_dev ? *h_dev : h_arg;  // the adaptive controller's step
  double v[3], W[3], R[9], pos[3], refpos[3], w[3];
  load3(b.linvel, i, v);
  load3(b.refangvel, i, W);
  // Obstacles at rest stay as built (the reference skips them,
  // Engine.cpp:532).  An obstacle given a velocity (a rotating drum or box)
  // moves with it kinematically: no force, no gravity.
  const bool kinematic = b.obstacle[i] != 0;
  if (kinematic && v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && W[0] == 0.0 && W[1] == 0.0 &&
      W[2] == 0.0)
    return;
  load3(b.pos, i, pos);
  load3(b.refpos, i, refpos);
  for (int q = 0; q < 9; ++q) R[q] = b.rot[9 * i + q];
  if (!kinematic && (!touched || touched[i])) {
    const double m = b.mass[i];
    double F[3], T[3], J[9], Ji[9];
    load3(force, i, F);
    load3(torque, i, T);
    for (int q = 0; q < 9; ++q) {
      J[q] = b.inertia[9 * i + q];
      Ji[q] = b.inverse[9 * i + q];
    }
    for (int q = 0; q < 3; ++q) v[q] += h * (F[q] / m);
    update_angular(W, R, J, Ji, T, h);
  }
  for (int q = 0; q < 3; ++q)
    if (grav.g[q] != 0.0 && !kinematic) v[q] += h * grav.g[q];
  for (int q = 0; q < 3; ++q) pos[q] += h * v[q];
  update_rotation(w, W, R, h);
  if (lane == 0) {
    for (int q = 0; q < 3; ++q) {
      b.linvel[3 * i + q] = v[q];
      b.refangvel[3 * i + q] = W[q];
      b.angvel[3 * i + q] = w[q];
      b.pos[3 * i + q] = pos[q];
    }
    for (int q = 0; q < 9; ++q) b.rot[9 * i + q] = R[q];
  }
  const int v0 = 3 * tri_offset[i], v1 = 3 * tri_offset[i + 1];
  // kIntegrateBatch replicas per lane at a time, their loads issued together
  for (int j0 = v0 + lane; j0 < v1; j0 += 64 * kIntegrateBatch) {
    double hx[kIntegrateBatch], hy[kIntegrateBatch], hz[kIntegrateBatch];
#pragma unroll
    for (int k = 0; k < kIntegrateBatch; ++k) {
      const int j = j0 + 64 * k;
      const bool in = j < v1;
      hx[k] = in ? rx[j] : 0.0;
      hy[k] = in ? ry[j] : 0.0;
      hz[k] = in ? rz[j] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kIntegrateBatch; ++k) {
      const int j = j0 + 64 * k;
      if (j < v1) update_vertex(x[j], y[j], z[j], hx[k], hy[k], hz[k], R, pos, refpos);
    }
  }
}


struct ForceScratch {
  DevBuf pf, incid, start, fill, entries, wide;
};
static ForceScratch g_fs;

int scan_counts(const int* counts, int n, long long* offs, hipStream_t st);

// Buffers of contacts_forces_impl for up to npairs pairs (before a capture).
int forces_reserve(int n, int npairs) {
  int rc;
  if ((rc = g_fs.pf.grow(sizeof(double) * 12 * (npairs + 1)))) return rc;
  if ((rc = g_fs.incid.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_fs.fill.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_fs.start.grow(sizeof(long long) * (n + 1)))) return rc;
  if ((rc = g_fs.entries.grow(sizeof(int) * (2 * static_cast<size_t>(npairs) + 1)))) return rc;
  if ((rc = g_fs.wide.grow(sizeof(int) * (static_cast<size_t>(n) + 2)))) return rc;
  return DH_OK;
}

// Per-pair force sums then per-particle sums in ascending partner order.
// With np_dev, npairs is a capacity and *np_dev the live count; no host
// synchronisation and no memset (capturable).
int contacts_forces_impl(const dh_bodies* bodies, const int* pairs_dev, int npairs,
                         const long long* np_dev, const dh_contact* contacts_dev,
                         long long contact_cap, const long long* pair_offset_dev,
                         const dh_force_params* fp, double* force_dev, double* torque_dev,
                         int* touched_dev, hipStream_t st) {
  const int n = bodies->n;
  int rc;
  if ((rc = forces_reserve(n, npairs))) return rc;
  const BodiesDev b = to_dev(bodies);
  const ForceConsts k{fp->spring, fp->damper, fp->friction};
  if (n <= kForcesSmallN) {  // one workgroup, one launch
    hipLaunchKernelGGL(k_forces_small, 1, 1024, 0, st, b, pairs_dev, npairs, np_dev,
                       contacts_dev, pair_offset_dev, k, contact_cap, g_fs.pf.as<double>(),
                       g_fs.entries.as<int>(), force_dev, torque_dev, touched_dev);
    DH_HIP_CHECK(hipGetLastError());
    return DH_OK;
  }
  zero_async(ZeroRegions{{g_fs.incid.as<unsigned>(), g_fs.fill.as<unsigned>(),
                          g_fs.wide.as<unsigned>(), nullptr},
                         {n + 1, n + 1, 1, 0}},
             st);
  if (npairs > 0) {
    hipLaunchKernelGGL(k_pair_forces, grid_for(npairs, 256), 256, 0, st, b, pairs_dev, npairs,
                       contacts_dev, pair_offset_dev, k, g_fs.pf.as<double>(),
                       g_fs.incid.as<int>(), np_dev, contact_cap);
    DH_HIP_CHECK(hipGetLastError());
  }
  if ((rc = scan_counts(g_fs.incid.as<int>(), n, g_fs.start.as<long long>(), st))) return rc;
  if (npairs > 0) {
    hipLaunchKernelGGL(k_fill_incidence, grid_for(npairs, 256), 256, 0, st, pairs_dev, npairs,
                       pair_offset_dev, g_fs.start.as<long long>(), g_fs.fill.as<int>(),
                       g_fs.entries.as<int>(), np_dev, contact_cap);
    DH_HIP_CHECK(hipGetLastError());
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_particle_sum, grid_for(n, 256), 256, 0, st, n, pairs_dev,
                       g_fs.start.as<long long>(), g_fs.incid.as<int>(), g_fs.entries.as<int>(),
                       g_fs.pf.as<double>(), force_dev, torque_dev, touched_dev,
                       g_fs.wide.as<int>());
    // a workgroup per listed particle, at most 1,024 workgroups
    hipLaunchKernelGGL(k_particle_sum_wide, static_cast<unsigned>(n < 1024 ? n : 1024), 256, 0,
                       st, g_fs.wide.as<int>(), pairs_dev, g_fs.start.as<long long>(),
                       g_fs.incid.as<int>(), g_fs.entries.as<int>(), g_fs.pf.as<double>(),
                       force_dev, torque_dev, touched_dev);
    DH_HIP_CHECK(hipGetLastError());
  }
  return DH_OK;
}

// dt_dev (may be null): the step is read there when the kernel runs.
int integrate_impl(const dh_bodies* bodies, const double* force_dev, const double* torque_dev,
                   const int* touched_dev, double dt, const double* dt_dev,
                   const double gravity[3], double* x_dev, double* y_dev, double* z_dev,
                   const double* refx_dev, const double* refy_dev, const double* refz_dev,
                   const int* tri_offset_dev, hipStream_t st) {
  if (bodies->n == 0) return DH_OK;
  Gravity g{{gravity[0], gravity[1], gravity[2]}};
  hipLaunchKernelGGL(k_integrate, bodies->n, 64, 0, st, to_dev(bodies), force_dev,
                     torque_dev, touched_dev, dt, dt_dev, g, x_dev, y_dev, z_dev, refx_dev,
                     refy_dev, refz_dev, tri_offset_dev);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}
}  // namespace dh

using namespace dh;

extern "C" {

int dh_contacts_forces(const dh_bodies* bodies, const int* pairs_dev, int npairs,
                       const dh_contact* contacts_dev, const int64_t* pair_offset_dev,
                       const dh_force_params* fp, double* force_dev, double* torque_dev,
                       int* touched_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bodies || !fp || !force_dev || !torque_dev || npairs < 0 ||
      (npairs > 0 && (!pairs_dev || !pair_offset_dev)))
    return set_error(DH_ERR_ARG, "dh_contacts_forces: bad arguments");
  return contacts_forces_impl(bodies, pairs_dev, npairs, nullptr, contacts_dev,
                              0x7fffffffffffffffLL,
                              reinterpret_cast<const long long*>(pair_offset_dev), fp, force_dev,
                              torque_dev, touched_dev, as_stream(stream));
}


int dh_integrate(dh_bodies* bodies, const double* force_dev, const double* torque_dev,
                 const int* touched_dev, double dt, const double gravity[3], double* x_dev,
                 double* y_dev, double* z_dev, const double* refx_dev, const double* refy_dev,
                 const double* refz_dev, const int* tri_offset_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bodies || !force_dev || !torque_dev || !gravity || !tri_offset_dev)
    return set_error(DH_ERR_ARG, "dh_integrate: bad arguments");
  return integrate_impl(bodies, force_dev, torque_dev, touched_dev, dt, nullptr, gravity, x_dev,
                        y_dev, z_dev, refx_dev, refy_dev, refz_dev, tri_offset_dev,
                        as_stream(stream));
}

// The single-particle forms run on the host: one particle's update is ~100
// floating-point operations, far below a kernel launch and two copies, and
// the host's libm sin/cos are the reference's (the device ones can differ in
// the last bit where expmap leaves its series, |h W| >= 0.1745).
int dh_update_angular(double refAngular[3], const double rotation[9], const double inertia[9],
                      const double inverse[9], const double torque[3], double step) {
  if (!refAngular || !rotation || !inertia || !inverse || !torque)
    return set_error(DH_ERR_ARG, "dh_update_angular: null pointer");
  update_angular(refAngular, rotation, inertia, inverse, torque, step);
  return DH_OK;
}

int dh_update_rotation_matrix(double angular[3], const double refAngular[3], double rotation[9],
                              double step) {
  if (!angular || !refAngular || !rotation)
    return set_error(DH_ERR_ARG, "dh_update_rotation_matrix: null pointer");
  update_rotation(angular, refAngular, rotation, step);
  return DH_OK;
}

}  // extern "C"
