// broadphase.hip — candidate particle pairs from a uniform hash grid.
//
// The reference has no broad phase on its engine path (Engine.cpp:272-295 is
// an O(N^2) loop with an optional bounding-sphere test; Peano supplies a grid
// in the app).  Here the candidate set is defined exactly: every pair (i<j),
// not both obstacles, whose particle AABBs inflated by their margins (plus a
// 0.5e-9 m band each) overlap.  The grid is a search structure only: cell
// size >= the largest inflated extent of a non-obstacle particle guarantees
// that overlapping boxes sit in neighbouring cells, so the grid result equals
// the all-pairs test.  "Large" particles -- obstacles (floors, funnels, boxes)
// and any particle whose inflated box outgrows the cell, which the reference's
// non-orthogonal rotation update (dynamics.cpp:78) can cause -- are kept out
// of the grid and tested against every particle.  Output pairs are sorted (i ascending, then
// j ascending) regardless of hash-bucket order.
#include <mutex>
#include <vector>

#include "runtime.hpp"

namespace dh {

struct Aabb {
  double lo[3], hi[3];
};

// Vertex replicas per lane that the box pass loads at once and keeps for the
// sphere pass (a particle of up to 64 * kAabbHeld replicas is read once).
constexpr int kAabbHeld = 8;

__global__ __launch_bounds__(64) void k_particle_aabb(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const int* __restrict__ tri_offset,
                                                      const double* __restrict__ eps,
                                                      Aabb* __restrict__ box,
                                                      Aabb* __restrict__ cbox,
                                                      double* __restrict__ sph) {
  const int i = blockIdx.x;
  const int lane = static_cast<int>(threadIdx.x);
  const int t0 = tri_offset[i], t1 = tri_offset[i + 1];
  // a particle with more triangles than one row block keeps the box of each
  // chunk of 64 of its triangles, chunk c at cbox[t0 / 64 + i + c] (an index
  // that grows with i and c, so no two chunks share it)
  const bool chunked = cbox && t1 - t0 > kRowBlock;
  const int v0 = 3 * t0, nv = t1 > t0 ? 3 * (t1 - t0) : 0;
  double plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
  // the held replicas (NaN past the end: min, max and the sphere skip NaN)
  double hx[kAabbHeld], hy[kAabbHeld], hz[kAabbHeld];
  auto load_block = [&](int b) {
#pragma unroll
    for (int k = 0; k < kAabbHeld; ++k) {
      const int v = b + 64 * k + lane;
      const bool in = v < nv;
      hx[k] = in ? x[v0 + v] : NAN;
      hy[k] = in ? y[v0 + v] : NAN;
      hz[k] = in ? z[v0 + v] : NAN;
    }
  };
  if (!chunked) {
    for (int b = 0; b < nv; b += 64 * kAabbHeld) {
      load_block(b);
#pragma unroll
      for (int k = 0; k < kAabbHeld; ++k) {
        plo[0] = fmin(plo[0], hx[k]);
        plo[1] = fmin(plo[1], hy[k]);
        plo[2] = fmin(plo[2], hz[k]);
        phi[0] = fmax(phi[0], hx[k]);
        phi[1] = fmax(phi[1], hy[k]);
        phi[2] = fmax(phi[2], hz[k]);
      }
    }
    for (int o = 32; o > 0; o >>= 1)
      for (int q = 0; q < 3; ++q) {
        plo[q] = fmin(plo[q], __shfl_xor(plo[q], o));
        phi[q] = fmax(phi[q], __shfl_xor(phi[q], o));
      }
  } else {
    // kChunkGroup chunks at a time, their loads issued together
    constexpr int kChunkGroup = 4;
    for (int g0 = t0; g0 < t1; g0 += 64 * kChunkGroup) {
      double lo[kChunkGroup][3], hi[kChunkGroup][3];
#pragma unroll
      for (int c = 0; c < kChunkGroup; ++c) {
        const int b0 = g0 + 64 * c, b1 = b0 + 64 < t1 ? b0 + 64 : t1;
        double cx[3], cy[3], cz[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int v = 3 * b0 + 64 * k + lane;
          const bool in = v < 3 * b1;
          cx[k] = in ? x[v] : NAN;
          cy[k] = in ? y[v] : NAN;
          cz[k] = in ? z[v] : NAN;
        }
        // from +-inf, as the box of no (or only NaN) coordinates
        lo[c][0] = fmin(fmin(fmin(INFINITY, cx[0]), cx[1]), cx[2]);
        lo[c][1] = fmin(fmin(fmin(INFINITY, cy[0]), cy[1]), cy[2]);
        lo[c][2] = fmin(fmin(fmin(INFINITY, cz[0]), cz[1]), cz[2]);
        hi[c][0] = fmax(fmax(fmax(-INFINITY, cx[0]), cx[1]), cx[2]);
        hi[c][1] = fmax(fmax(fmax(-INFINITY, cy[0]), cy[1]), cy[2]);
        hi[c][2] = fmax(fmax(fmax(-INFINITY, cz[0]), cz[1]), cz[2]);
      }
      for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int c = 0; c < kChunkGroup; ++c)
          for (int q = 0; q < 3; ++q) {
            lo[c][q] = fmin(lo[c][q], __shfl_xor(lo[c][q], o));
            hi[c][q] = fmax(hi[c][q], __shfl_xor(hi[c][q], o));
          }
#pragma unroll
      for (int c = 0; c < kChunkGroup; ++c) {
        const int b0 = g0 + 64 * c;
        if (b0 >= t1) break;
        if (lane == 0) {
          Aabb bx;
          for (int q = 0; q < 3; ++q) {
            bx.lo[q] = lo[c][q];
            bx.hi[q] = hi[c][q];
          }
          cbox[t0 / 64 + i + (b0 - t0) / 64] = bx;
        }
        for (int q = 0; q < 3; ++q) {
          plo[q] = fmin(plo[q], lo[c][q]);
          phi[q] = fmax(phi[q], hi[c][q]);
        }
      }
    }
  }
  if (lane == 0) {
    const double m = eps[i] + 0.5e-9;
    Aabb bx;
    for (int q = 0; q < 3; ++q) {
      bx.lo[q] = plo[q] - m;
      bx.hi[q] = phi[q] + m;
    }
    box[i] = bx;
  }
  if (sph) {
    // a bounding sphere: the vertex box's centre and the largest vertex
    // distance from it (NaN coordinates drop out, as from the box); the
    // replicas come from registers when one block held them all
    const double c[3] = {0.5 * (plo[0] + phi[0]), 0.5 * (plo[1] + phi[1]),
                         0.5 * (plo[2] + phi[2])};
    double r2 = 0.0;
    const bool held = !chunked && nv <= 64 * kAabbHeld;
    for (int b = 0; b < nv; b += 64 * kAabbHeld) {
      if (!held) load_block(b);
#pragma unroll
      for (int k = 0; k < kAabbHeld; ++k) {
        const double dx = hx[k] - c[0], dy = hy[k] - c[1], dz = hz[k] - c[2];
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
      }
    }
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o));
    if (lane == 0) {
      sph[4 * i] = c[0];
      sph[4 * i + 1] = c[1];
      sph[4 * i + 2] = c[2];
      sph[4 * i + 3] = sqrt(r2);
    }
  }
}

__device__ __forceinline__ bool overlap(const Aabb& a, const Aabb& b) {
  for (int q = 0; q < 3; ++q)
    if (a.lo[q] > b.hi[q] || b.lo[q] > a.hi[q]) return false;
  return true;
}

__device__ __forceinline__ void cell_of(const Aabb& b, double inv_h, long long c[3]) {
  for (int q = 0; q < 3; ++q) c[q] = static_cast<long long>(floor(0.5 * (b.lo[q] + b.hi[q]) * inv_h));
}

__device__ __forceinline__ unsigned hash_cell(long long cx, long long cy, long long cz,
                                              unsigned mask) {
  unsigned long long h = static_cast<unsigned long long>(cx) * 0x9E3779B97F4A7C15ull ^
                         static_cast<unsigned long long>(cy) * 0xC2B2AE3D27D4EB4Full ^
                         static_cast<unsigned long long>(cz) * 0x165667B19E3779F9ull;
  h ^= h >> 29;
  return static_cast<unsigned>(h) & mask;
}

// The grid is a cell list sorted by hash bucket: bucket b holds the
// particles sorted[cstart[b] .. cstart[b+1]) (a counting sort: counts,
// exclusive scan, scatter), so a bucket is one contiguous range and its
// members load without pointer chasing.
struct GridArgs {
  const Aabb* box;
  const int* obstacle;  // may be null
  int n;
  double h, inv_h;
  unsigned mask;             // mask + 1 buckets
  int* ccount;               // particles per bucket
  int* cfill;                // scatter cursor per bucket
  const long long* cstart;   // bucket starts (mask + 2 entries)
  int* sorted;               // particle ids by bucket
  unsigned* cellh;           // bucket of each particle (large: unused)
  int* large;                // indices of large particles (any order)
  int* nlarge;
  // partners of a large particle are searched in chunks of `chunk`
  // candidates j (nchunks per particle, chunk c covers [c*chunk, (c+1)*chunk)),
  // one wave per (large particle, chunk); lcount[slot * nchunks + c] holds
  // each chunk's partner count
  int chunk, nchunks;
  int* lcount;
  // the first kBpKeep partners of each particle as the count pass found
  // them (discovery order), so that the fill pass only sorts them
  int* keep;
};

__device__ __forceinline__ bool is_large(const GridArgs& g, int i, const Aabb& b) {
  if (g.obstacle && g.obstacle[i]) return true;
  for (int q = 0; q < 3; ++q)
    if (!(b.hi[q] - b.lo[q] <= g.h)) return true;
  return false;
}

__global__ __launch_bounds__(256) void k_grid_bin(GridArgs g) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= g.n) return;
  const Aabb b = g.box[i];
  if (is_large(g, i, b)) {
    g.large[atomicAdd(g.nlarge, 1)] = i;
    return;
  }
  long long c[3];
  cell_of(b, g.inv_h, c);
  const unsigned hsh = hash_cell(c[0], c[1], c[2], g.mask);
  g.cellh[i] = hsh;
  atomicAdd(&g.ccount[hsh], 1);
}

__global__ __launch_bounds__(256) void k_grid_scatter(GridArgs g) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= g.n) return;
  if (is_large(g, i, g.box[i])) return;
  const unsigned hsh = g.cellh[i];
  g.sorted[g.cstart[hsh] + atomicAdd(&g.cfill[hsh], 1)] = i;
}

// One wave searches the partners j > i of one particle i: the 27 neighbour
// buckets are walked by 27 lanes at once, one member per lane per round, and
// the large-particle list is strided over the wave.  A large particle's scan
// over all j is split into chunks, one wave each.  Finds are compacted with ballots, so every lane sees the
// running count.
constexpr int kBpWaves = 4;    // particles per workgroup
constexpr int kBpList = 512;   // partners sorted in LDS per particle (else a serial sort)
constexpr int kBpKeep = 64;    // partners kept by the count pass per particle (one per lane)

__device__ __forceinline__ int bp_lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
}

// Calls put(position, j) on the finding lane for every partner of a
// particle i that is not large, positions 0.. in discovery order; returns
// the number of partners (wave-uniform).
template <typename F>
__device__ int grid_partners(const GridArgs& g, int i, const Aabb& bi, int lane, F put) {
  long long c[3];
  cell_of(bi, g.inv_h, c);
  const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
  const unsigned hsh = lane < 27 ? hash_cell(c[0] + dx, c[1] + dy, c[2] + dz, g.mask) : 0u;
  // distinct neighbour cells may share a bucket: its lowest lane walks it
  bool walk = lane < 27;
  for (int l = 0; l < 26; ++l) {
    const unsigned o = __shfl(hsh, l);
    if (l < lane && o == hsh) walk = false;
  }
  long long pos = walk ? g.cstart[hsh] : 0, end = walk ? g.cstart[hsh + 1] : 0;
  int found = 0;
  while (__ballot(pos < end)) {
    bool e = false;
    int cand = -1;
    if (pos < end) {
      cand = g.sorted[pos];
      if (cand > i) {
        const Aabb bj = g.box[cand];
        long long cj[3];
        cell_of(bj, g.inv_h, cj);
        e = cj[0] - c[0] >= -1 && cj[0] - c[0] <= 1 && cj[1] - c[1] >= -1 && cj[1] - c[1] <= 1 &&
            cj[2] - c[2] >= -1 && cj[2] - c[2] <= 1 && overlap(bi, bj);
      }
      ++pos;
    }
    const unsigned long long mask = __ballot(e);
    if (e) put(found + bp_lane_prefix(mask), cand);
    found += __popcll(mask);
  }
  const int nl = *g.nlarge;
  for (int o0 = 0; o0 < nl; o0 += 64) {
    const int o = o0 + lane;
    bool e = false;
    int cand = -1;
    if (o < nl) {
      cand = g.large[o];
      e = cand > i && overlap(bi, g.box[cand]);
    }
    const unsigned long long mask = __ballot(e);
    if (e) put(found + bp_lane_prefix(mask), cand);
    found += __popcll(mask);
  }
  return found;
}

// Partners j of a large particle i in chunk c: every j > i of the chunk
// (not both obstacles), in ascending order; put(position, j) as above.
template <typename F>
__device__ int large_chunk(const GridArgs& g, int i, const Aabb& bi, int c, int lane, F put) {
  const bool obst_i = g.obstacle && g.obstacle[i];
  const int lo = c * g.chunk > i + 1 ? c * g.chunk : i + 1;
  const int hi = (c + 1) * g.chunk < g.n ? (c + 1) * g.chunk : g.n;
  int found = 0;
  // four strips of 64 per round: their loads are issued together
  for (int j0 = lo; j0 < hi; j0 += 256) {
    bool e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + 64 * q + lane;
      e[q] = j < hi && !(obst_i && g.obstacle && g.obstacle[j]) && overlap(bi, g.box[j]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const unsigned long long mask = __ballot(e[q]);
      if (e[q]) put(found + bp_lane_prefix(mask), j0 + 64 * q + lane);
      found += __popcll(mask);
    }
  }
  return found;
}

// Large particles, one wave per (slot, chunk) item, strided over a
// persistent grid (the number of large particles is known on the device
// only): chunk partner counts.
__global__ __launch_bounds__(64 * kBpWaves) void k_large_count(GridArgs g) {
  const int lane = threadIdx.x & 63;
  const long long nw = static_cast<long long>(gridDim.x) * kBpWaves;
  const long long items = static_cast<long long>(*g.nlarge) * g.nchunks;
  auto none = [](int, int) {};
  for (long long t = static_cast<long long>(blockIdx.x) * kBpWaves + (threadIdx.x >> 6);
       t < items; t += nw) {
    const int o = static_cast<int>(t / g.nchunks), c = static_cast<int>(t - o * static_cast<long long>(g.nchunks));
    const int i = g.large[o];
    const int m = large_chunk(g, i, g.box[i], c, lane, none);
    if (lane == 0) g.lcount[t] = m;
  }
}

// Partner count of each large particle: the sum of its chunk counts.
__global__ __launch_bounds__(64 * kBpWaves) void k_large_sum(GridArgs g, int* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * kBpWaves;
  const int nl = *g.nlarge;
  for (int o = blockIdx.x * kBpWaves + (threadIdx.x >> 6); o < nl; o += nw) {
    int m = 0;
    for (int c = lane; c < g.nchunks; c += 64) m += g.lcount[static_cast<long long>(o) * g.nchunks + c];
    for (int k = 32; k > 0; k >>= 1) m += __shfl_xor(m, k);
    if (lane == 0) count[g.large[o]] = m;
  }
}

// Writes each chunk's partners after those of the particle's earlier chunks.
__global__ __launch_bounds__(64 * kBpWaves) void k_large_fill(GridArgs g,
                                                               const long long* __restrict__ start,
                                                               long long cap,
                                                               int* __restrict__ pairs) {
  const int lane = threadIdx.x & 63;
  const long long nw = static_cast<long long>(gridDim.x) * kBpWaves;
  const long long items = static_cast<long long>(*g.nlarge) * g.nchunks;
  for (long long t = static_cast<long long>(blockIdx.x) * kBpWaves + (threadIdx.x >> 6);
       t < items; t += nw) {
    const int o = static_cast<int>(t / g.nchunks), c = static_cast<int>(t - o * static_cast<long long>(g.nchunks));
    const int i = g.large[o];
    long long before = 0;
    for (int q = lane; q < c; q += 64) before += g.lcount[t - c + q];
    for (int k = 32; k > 0; k >>= 1) before += __shfl_xor(before, k);
    const long long s = start[i] + before;
    large_chunk(g, i, g.box[i], c, lane, [&](int pos, int j) {
      const long long q = s + pos;
      if (q < cap) {
        pairs[2 * q] = i;
        pairs[2 * q + 1] = j;
      }
    });
  }
}

// grid_partners for two particles per wave: half h = lane / 32 searches
// particle i (each half's i, bi and `found` are its own); the same partner
// set as grid_partners, in another discovery order.
template <typename F>
__device__ int grid_partners_half(const GridArgs& g, int i, const Aabb& bi, bool active, int lane,
                                  F put) {
  const int h = lane >> 5, sl = lane & 31;
  long long c[3];
  cell_of(bi, g.inv_h, c);
  const int dx = sl / 9 - 1, dy = (sl / 3) % 3 - 1, dz = sl % 3 - 1;
  const unsigned hsh = (active && sl < 27) ? hash_cell(c[0] + dx, c[1] + dy, c[2] + dz, g.mask) : 0u;
  bool walk = active && sl < 27;
  for (int l = 0; l < 26; ++l) {
    const unsigned o = __shfl(hsh, (h << 5) + l);
    if (l < sl && o == hsh) walk = false;
  }
  long long pos = walk ? g.cstart[hsh] : 0, end = walk ? g.cstart[hsh + 1] : 0;
  const unsigned below = (1u << sl) - 1u;
  int found = 0;
  while (__ballot(pos < end)) {
    bool e = false;
    int cand = -1;
    if (pos < end) {
      cand = g.sorted[pos];
      if (cand > i) {
        const Aabb bj = g.box[cand];
        long long cj[3];
        cell_of(bj, g.inv_h, cj);
        e = cj[0] - c[0] >= -1 && cj[0] - c[0] <= 1 && cj[1] - c[1] >= -1 && cj[1] - c[1] <= 1 &&
            cj[2] - c[2] >= -1 && cj[2] - c[2] <= 1 && overlap(bi, bj);
      }
      ++pos;
    }
    const unsigned hm = static_cast<unsigned>(__ballot(e) >> (32 * h));
    if (e) put(found + __popc(hm & below), cand);
    found += __popc(hm);
  }
  const int nl = *g.nlarge;
  for (int o0 = 0; o0 < nl; o0 += 32) {
    const int o = o0 + sl;
    bool e = false;
    int cand = -1;
    if (active && o < nl) {
      cand = g.large[o];
      e = cand > i && overlap(bi, g.box[cand]);
    }
    const unsigned hm = static_cast<unsigned>(__ballot(e) >> (32 * h));
    if (e) put(found + __popc(hm & below), cand);
    found += __popc(hm);
  }
  return found;
}

// Partner counts, two particles per wave (the walk is latency bound: two
// searches share each round of dependent loads).
__global__ __launch_bounds__(64 * kBpWaves) void k_grid_count2(GridArgs g,
                                                                int* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int i = (blockIdx.x * kBpWaves + (threadIdx.x >> 6)) * 2 + (lane >> 5);
  const bool valid = i < g.n;
  const Aabb bi = g.box[valid ? i : 0];
  const bool active = valid && !is_large(g, i, bi);
  int* kept = g.keep + static_cast<long long>(valid ? i : 0) * kBpKeep;
  auto keep = [&](int pos, int j) {
    if (pos < kBpKeep) kept[pos] = j;
  };
  const int m = grid_partners_half(g, i, bi, active, lane, keep);
  // large particles: k_large_count / k_large_sum, launched after this
  if (valid && (lane & 31) == 0) count[i] = active ? m : 0;
}

__global__ __launch_bounds__(64 * kBpWaves) void k_grid_count(GridArgs g,
                                                               int* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kBpWaves + (threadIdx.x >> 6);
  if (i >= g.n) return;  // wave-uniform
  const Aabb bi = g.box[i];
  int* kept = g.keep + static_cast<long long>(i) * kBpKeep;
  auto keep = [&](int pos, int j) {
    if (pos < kBpKeep) kept[pos] = j;
  };
  // large particles: k_large_count / k_large_sum, launched after this
  const int m = is_large(g, i, bi) ? 0 : grid_partners(g, i, bi, lane, keep);
  if (lane == 0) count[i] = m;
}

// Writes the partners of i at start[i].. in ascending order (pairs beyond
// cap are not written).
__global__ __launch_bounds__(64 * kBpWaves) void k_grid_fill(GridArgs g,
                                                              const long long* __restrict__ start,
                                                              long long cap,
                                                              int* __restrict__ pairs) {
  __shared__ int list[kBpWaves][kBpList];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = blockIdx.x * kBpWaves + w;
  if (i >= g.n) return;  // wave-uniform
  const long long s = start[i];
  const long long m = start[i + 1] - s;
  const Aabb bi = g.box[i];
  auto direct = [&](int pos, int j) {
    const long long q = s + pos;
    if (q < cap) {
      pairs[2 * q] = i;
      pairs[2 * q + 1] = j;
    }
  };
  if (is_large(g, i, bi)) return;  // k_large_fill
  if (m <= kBpKeep) {
    // the count pass kept them all: lane e ranks entry e among the m
    // (distinct) partners by shuffles and writes it in place
    const int v = lane < m ? g.keep[static_cast<long long>(i) * kBpKeep + lane] : 0;
    int rank = 0;
    for (int q = 0; q < m; ++q) {
      const int u = __shfl(v, q);
      rank += u < v ? 1 : 0;
    }
    if (lane < m) direct(rank, v);
    return;
  }
  if (m > kBpList) {
    // more partners than the LDS list holds: unsorted, then a serial sort
    grid_partners(g, i, bi, lane, direct);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    if (lane == 0) {
      const long long e = s + m < cap ? s + m : cap;
      for (long long a = s + 1; a < e; ++a) {
        const int v = pairs[2 * a + 1];
        long long b = a - 1;
        while (b >= s && pairs[2 * b + 1] > v) {
          pairs[2 * (b + 1) + 1] = pairs[2 * b + 1];
          --b;
        }
        pairs[2 * (b + 1) + 1] = v;
      }
    }
    return;
  }
  int* lst = list[w];
  const int found = grid_partners(g, i, bi, lane, [&](int pos, int j) {
    if (pos < kBpList) lst[pos] = j;
  });
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
  const int mm = found < kBpList ? found : kBpList;
  // partners are distinct: the rank of v is the number of smaller partners
  for (int e = lane; e < mm; e += 64) {
    const int v = lst[e];
    int rank = 0;
    for (int q = 0; q < mm; ++q) rank += lst[q] < v ? 1 : 0;
    direct(rank, v);
  }
}

// Scenes of at most kSmallScene particles: every pair (i < j) tested in one
// workgroup of 16 waves, in place of the grid's dozen launches.  The same
// candidate set (boxes overlap, not both obstacles) in the same order (i
// ascending, then j ascending), with the same outputs: start[0..n] (start[n]
// the total, which may exceed cap) and the pairs that fit cap.  A wave per
// particle i sweeps j > i in strips of 64: counts, their prefix sums, then
// the same sweep writes each partner at start[i] + its rank.
constexpr int kSmallScene = 1024;

__global__ __launch_bounds__(1024) void k_small_pairs(const Aabb* __restrict__ box,
                                                      const int* __restrict__ obstacle, int n,
                                                      long long* __restrict__ start,
                                                      long long cap, int* __restrict__ pairs) {
  // the boxes and obstacle flags staged once: both sweeps read LDS only
  __shared__ Aabb s_box[kSmallScene];
  __shared__ unsigned char s_obst[kSmallScene];
  __shared__ long long s_start[kSmallScene + 1];
  __shared__ long long s_wave[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = threadIdx.x;
  if (t < n) {
    s_box[t] = box[t];
    s_obst[t] = obstacle && obstacle[t] ? 1 : 0;
  }
  __syncthreads();
  for (int i = w; i < n; i += 16) {
    const Aabb bi = s_box[i];
    const bool oi = s_obst[i] != 0;
    int m = 0;
    for (int j0 = i + 1; j0 < n; j0 += 64) {
      const int j = j0 + lane;
      const bool e = j < n && !(oi && s_obst[j]) && overlap(bi, s_box[j]);
      m += __popcll(__ballot(e));
    }
    if (lane == 0) s_start[i] = m;
  }
  __syncthreads();
  // exclusive prefix sums: one count per thread, waves, then wave totals
  const long long c = t < n ? s_start[t] : 0;
  long long incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const long long u = __shfl_up(incl, o);
    incl += lane >= o ? u : 0;
  }
  if (lane == 63) s_wave[w] = incl;
  __syncthreads();
  long long before = 0;
  for (int q = 0; q < w; ++q) before += s_wave[q];
  const long long excl = before + incl - c;
  if (t < n) {
    s_start[t] = excl;
    start[t] = excl;
  }
  if (t == (n > 0 ? n - 1 : 0)) {
    s_start[n] = n > 0 ? excl + c : 0;
    start[n] = n > 0 ? excl + c : 0;
  }
  __syncthreads();
  for (int i = w; i < n; i += 16) {
    const Aabb bi = s_box[i];
    const bool oi = s_obst[i] != 0;
    const long long s0 = s_start[i];
    int found = 0;
    for (int j0 = i + 1; j0 < n; j0 += 64) {
      const int j = j0 + lane;
      const bool e = j < n && !(oi && s_obst[j]) && overlap(bi, s_box[j]);
      const unsigned long long mask = __ballot(e);
      const long long q = s0 + found + bp_lane_prefix(mask);
      if (e && q < cap) {
        pairs[2 * q] = i;
        pairs[2 * q + 1] = j;
      }
      found += __popcll(mask);
    }
  }
}

int scan_counts(const int* counts, int n, long long* offs, hipStream_t st);
int scan_counts_with(const int* counts, int n, long long* offs, long long* sums,
                     hipStream_t st);
int scan_sums_len(int n);
int scan_reserve(int n);

struct BpScratch {
  DevBuf box, ccount, cfill, cstart, sorted, cellh, count, start, err, obst, lcount, keep, cbox,
      sph;
};
static BpScratch g_bp;

// Candidates per large-particle chunk: 1024 (four rounds of 256 loads per
// wave), more for very large n so that a particle has at most 256 chunks.
static int large_chunk_len(int n) {
  int c = 1024;
  while (static_cast<long long>(c) * 256 < n) c *= 2;
  return c;
}

static int large_chunks(int n) {
  const int c = large_chunk_len(n);
  return (n + c - 1) / c > 0 ? (n + c - 1) / c : 1;
}

// Persistent grid of the large-particle kernels.
static unsigned large_grid(long long items) {
  const long long g = (items + kBpWaves - 1) / kBpWaves;
  return static_cast<unsigned>(g < 1 ? 1 : (g > 2048 ? 2048 : g));
}

static unsigned table_size(int n) {
  unsigned tbl = 1;
  while (tbl < 2u * static_cast<unsigned>(n)) tbl <<= 1;
  return tbl;
}

// The device buffers of one broad phase over n particles and ntri
// triangles (entries of each, in order of the layout below).
struct BpBuffers {
  Aabb* box;           // n + 1
  Aabb* cbox;          // ntri / 64 + n + 1
  int* ccount;         // tbl + 1
  int* cfill;          // tbl
  long long* cstart;   // tbl + 1
  int* sorted;         // n + 1
  unsigned* cellh;     // n + 1
  int* count;          // n + 1
  long long* start;    // n + 1
  int* nlarge;         // 1
  int* large;          // n + 1
  int* lcount;         // (n + 1) * large_chunks(n)
  int* keep;           // (n + 1) * kBpKeep
  long long* sums;     // scan_sums_len(max(tbl, n)); null: the scan's own scratch
  double* sph;         // 4 (n + 1): bounding spheres (centre, radius); may be null
};

static BpBuffers library_buffers() {
  return BpBuffers{g_bp.box.as<Aabb>(),     g_bp.cbox.as<Aabb>(),    g_bp.ccount.as<int>(),
                   g_bp.cfill.as<int>(),    g_bp.cstart.as<long long>(), g_bp.sorted.as<int>(),
                   g_bp.cellh.as<unsigned>(), g_bp.count.as<int>(),  g_bp.start.as<long long>(),
                   g_bp.err.as<int>(),      g_bp.obst.as<int>(),     g_bp.lcount.as<int>(),
                   g_bp.keep.as<int>(),     nullptr,                 g_bp.sph.as<double>()};
}

static int scan_into(const BpBuffers& b, const int* counts, int n, long long* offs,
                     hipStream_t st) {
  return b.sums ? scan_counts_with(counts, n, offs, b.sums, st) : scan_counts(counts, n, offs, st);
}

// The particle boxes of the last broad phase (lo xyz, hi xyz per particle:
// the vertex box inflated by the margin and 0.5e-9 m), valid until the next
// call on more particles.
const double* broadphase_boxes() { return g_bp.box.as<double>(); }
const double* broadphase_block_boxes() { return g_bp.cbox.as<double>(); }
// The particles' bounding spheres (x, y, z, r) of the last broad phase.
const double* broadphase_spheres() { return g_bp.sph.as<double>(); }

// The same boxes for the current geometry, without the grid (for callers
// of the detection that did not run the broad phase).
int particle_boxes(const dh_geometry* geom, const double** out, hipStream_t st) {
  const int n = geom->nparticles;
  int rc;
  if ((rc = g_bp.box.grow(sizeof(Aabb) * (n + 1)))) return rc;
  if ((rc = g_bp.cbox.grow(sizeof(Aabb) * ((geom->ntri_total + 63) / 64 + n + 1)))) return rc;
  if ((rc = g_bp.sph.grow(sizeof(double) * 4 * (n + 1)))) return rc;
  if (n > 0)
    hipLaunchKernelGGL(k_particle_aabb, n, 64, 0, st, geom->x, geom->y, geom->z,
                       geom->tri_offset, geom->eps, g_bp.box.as<Aabb>(), g_bp.cbox.as<Aabb>(),
                       g_bp.sph.as<double>());
  DH_HIP_CHECK(hipGetLastError());
  *out = g_bp.box.as<double>();
  return DH_OK;
}

// Buffers of the broad phase for n particles (call before a capture).
int broadphase_reserve(int n, int ntri) {
  int rc;
  const unsigned tbl = table_size(n);
  if ((rc = g_bp.box.grow(sizeof(Aabb) * (n + 1)))) return rc;
  if ((rc = g_bp.cbox.grow(sizeof(Aabb) * ((ntri + 63) / 64 + n + 1)))) return rc;
  if ((rc = g_bp.sph.grow(sizeof(double) * 4 * (n + 1)))) return rc;
  if ((rc = g_bp.ccount.grow(sizeof(int) * (tbl + 1)))) return rc;
  if ((rc = g_bp.cfill.grow(sizeof(int) * tbl))) return rc;
  if ((rc = g_bp.cstart.grow(sizeof(long long) * (tbl + 1)))) return rc;
  if ((rc = g_bp.sorted.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_bp.cellh.grow(sizeof(unsigned) * (n + 1)))) return rc;
  if ((rc = g_bp.count.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_bp.start.grow(sizeof(long long) * (n + 1)))) return rc;
  if ((rc = g_bp.err.grow(sizeof(int)))) return rc;
  if ((rc = g_bp.obst.grow(sizeof(int) * (n + 1)))) return rc;
  // every particle may be large (a cell smaller than the particles)
  if ((rc = g_bp.lcount.grow(sizeof(int) * static_cast<size_t>(n + 1) * large_chunks(n))))
    return rc;
  if ((rc = g_bp.keep.grow(sizeof(int) * static_cast<size_t>(n + 1) * kBpKeep))) return rc;
  return scan_reserve(static_cast<int>(tbl) > n ? static_cast<int>(tbl) : n);
}

// Clears the counters the grid passes add to (fill kernels, no memset, so
// that captured work may contain it).
static void grid_clear(const BpBuffers& b, unsigned tbl, hipStream_t st) {
  zero_async(ZeroRegions{{reinterpret_cast<unsigned*>(b.nlarge),
                          reinterpret_cast<unsigned*>(b.ccount),
                          reinterpret_cast<unsigned*>(b.cfill), nullptr},
                         {1, static_cast<long long>(tbl), static_cast<long long>(tbl), 0}},
             st);
}

// Boxes, the sorted cell list and the partner counts, on cleared counters.
static int grid_build(const dh_geometry* geom, const int* obstacle_dev, double cell,
                      const BpBuffers& b, GridArgs& g, hipStream_t st) {
  const int n = geom->nparticles;
  const unsigned tbl = table_size(n);
  hipLaunchKernelGGL(k_particle_aabb, n, 64, 0, st, geom->x, geom->y, geom->z,
                     geom->tri_offset, geom->eps, b.box, b.cbox, b.sph);
  g.box = b.box;
  g.obstacle = obstacle_dev;
  g.n = n;
  g.h = cell;
  g.inv_h = 1.0 / cell;
  g.mask = tbl - 1;
  g.ccount = b.ccount;
  g.cfill = b.cfill;
  g.cstart = b.cstart;
  g.sorted = b.sorted;
  g.cellh = b.cellh;
  g.large = b.large;
  g.nlarge = b.nlarge;
  g.chunk = large_chunk_len(n);
  g.nchunks = large_chunks(n);
  g.lcount = b.lcount;
  g.keep = b.keep;
  hipLaunchKernelGGL(k_grid_bin, grid_for(n, 256), 256, 0, st, g);
  int rc;
  if ((rc = scan_into(b, b.ccount, static_cast<int>(tbl), b.cstart, st))) return rc;
  hipLaunchKernelGGL(k_grid_scatter, grid_for(n, 256), 256, 0, st, g);
  hipLaunchKernelGGL(k_grid_count2, grid_for(n, 2 * kBpWaves), 64 * kBpWaves, 0, st, g, b.count);
  const long long items = static_cast<long long>(n) * g.nchunks;  // bound on the device count
  hipLaunchKernelGGL(k_large_count, large_grid(items), 64 * kBpWaves, 0, st, g);
  hipLaunchKernelGGL(k_large_sum, large_grid(n), 64 * kBpWaves, 0, st, g, b.count);
  DH_HIP_CHECK(hipGetLastError());
  return scan_into(b, b.count, n, b.start, st);
}

static void grid_fill(const GridArgs& g, const BpBuffers& b, long long cap, int* pairs,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_grid_fill, grid_for(g.n, kBpWaves), 64 * kBpWaves, 0, st, g, b.start, cap,
                     pairs);
  hipLaunchKernelGGL(k_large_fill, large_grid(static_cast<long long>(g.n) * g.nchunks),
                     64 * kBpWaves, 0, st, g, b.start, cap, pairs);
}

// The eager broad phase on the buffers b: the total is read back, and the
// pairs are written only if it fits cap.
static int broadphase_run(const dh_geometry* geom, const int* obstacle_dev, double cell,
                          const BpBuffers& b, int* pairs, long long cap, long long* npairs,
                          hipStream_t st) {
  const int n = geom->nparticles;
  if (n <= kSmallScene) {
    // one workgroup tests every pair; nothing is written past cap, so the
    // pairs come with the count
    if (n > 0) {
      hipLaunchKernelGGL(k_particle_aabb, n, 64, 0, st, geom->x, geom->y, geom->z,
                         geom->tri_offset, geom->eps, b.box, b.cbox, b.sph);
      hipLaunchKernelGGL(k_small_pairs, 1, 1024, 0, st, b.box, obstacle_dev, n, b.start, cap,
                         pairs);
      DH_HIP_CHECK(hipGetLastError());
    }
    long long total = 0;
    if (n > 0)
      DH_HIP_CHECK(hipMemcpyAsync(&total, b.start + n, sizeof(long long), hipMemcpyDeviceToHost,
                                  st));
    DH_HIP_CHECK(hipStreamSynchronize(st));
    *npairs = total;
    return total > cap ? DH_ERR_CAPACITY : DH_OK;
  }
  grid_clear(b, table_size(n), st);
  GridArgs g;
  int rc;
  if ((rc = grid_build(geom, obstacle_dev, cell, b, g, st))) return rc;
  long long total = 0;
  DH_HIP_CHECK(hipMemcpyAsync(&total, b.start + n, sizeof(long long), hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  *npairs = total;
  if (*npairs > cap) return DH_ERR_CAPACITY;
  grid_fill(g, b, cap, pairs, st);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int broadphase_impl(const dh_geometry* geom, const int* obstacle_dev, double cell, int* pairs,
                    long long cap, long long* npairs, hipStream_t st) {
  int rc;
  if ((rc = broadphase_reserve(geom->nparticles, geom->ntri_total))) return rc;
  return broadphase_run(geom, obstacle_dev, cell, library_buffers(), pairs, cap, npairs, st);
}

// The broad phase for a captured step: the same kernels with no host
// read-back.  *np_dev (the total, which may exceed cap) is left at the
// returned device address; pairs beyond cap are not written.
int broadphase_graph(const dh_geometry* geom, const int* obstacle_dev, double cell, int* pairs,
                     long long cap, const long long** np_dev, hipStream_t st) {
  const int n = geom->nparticles;
  int rc;
  if ((rc = broadphase_reserve(n, geom->ntri_total))) return rc;
  const BpBuffers b = library_buffers();
  *np_dev = b.start + n;
  if (n <= kSmallScene) {
    if (n > 0)
      hipLaunchKernelGGL(k_particle_aabb, n, 64, 0, st, geom->x, geom->y, geom->z,
                         geom->tri_offset, geom->eps, b.box, b.cbox, b.sph);
    hipLaunchKernelGGL(k_small_pairs, 1, 1024, 0, st, b.box, obstacle_dev, n, b.start, cap,
                       pairs);
    DH_HIP_CHECK(hipGetLastError());
    return DH_OK;
  }
  grid_clear(b, table_size(n), st);
  GridArgs g;
  if ((rc = grid_build(geom, obstacle_dev, cell, b, g, st))) return rc;
  grid_fill(g, b, cap, pairs, st);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

// ------------------------------------------------ broad phase, no allocation
// One statically allocated device array holds every buffer of a broad phase
// of at most DH_NOALLOC_BP_PARTICLES particles and DH_NOALLOC_BP_TRIANGLES
// triangles (8,523,264 bytes).  Each device keeps its own copy for the life of
// the process; every call on a device shares it, so calls that overlap on one
// device race.
constexpr size_t kNaAlign = 256;

static size_t na_round(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }

// Carves the buffers of (n, ntri) out of base; returns the bytes used.
static size_t noalloc_layout(char* base, int n, int ntri, BpBuffers* b) {
  const size_t tbl = table_size(n);
  const size_t np1 = static_cast<size_t>(n) + 1;
  const int scan_n = static_cast<int>(tbl) > n ? static_cast<int>(tbl) : n;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += na_round(bytes);
    return p;
  };
  BpBuffers t;
  t.box = reinterpret_cast<Aabb*>(take(sizeof(Aabb) * np1));
  t.cbox = reinterpret_cast<Aabb*>(take(sizeof(Aabb) * ((ntri + 63) / 64 + np1)));
  t.ccount = reinterpret_cast<int*>(take(sizeof(int) * (tbl + 1)));
  t.cfill = reinterpret_cast<int*>(take(sizeof(int) * tbl));
  t.cstart = reinterpret_cast<long long*>(take(sizeof(long long) * (tbl + 1)));
  t.sorted = reinterpret_cast<int*>(take(sizeof(int) * np1));
  t.cellh = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * np1));
  t.count = reinterpret_cast<int*>(take(sizeof(int) * np1));
  t.start = reinterpret_cast<long long*>(take(sizeof(long long) * np1));
  t.nlarge = reinterpret_cast<int*>(take(sizeof(int)));
  t.large = reinterpret_cast<int*>(take(sizeof(int) * np1));
  t.lcount = reinterpret_cast<int*>(take(sizeof(int) * np1 * large_chunks(n)));
  t.keep = reinterpret_cast<int*>(take(sizeof(int) * np1 * kBpKeep));
  t.sums = reinterpret_cast<long long*>(take(sizeof(long long) * scan_sums_len(scan_n)));
  t.sph = nullptr;  // the detection's sphere test is not part of this entry point
  if (b) *b = t;
  return off;
}

// Bytes of the largest layout (every buffer grows with n and ntri).
constexpr size_t kNaScratchBytes = 8523264;

__device__ __align__(256) char scratch[kNaScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_na_ptr[kMaxDevices];
static std::mutex g_na_mu;

// The device address of this device's `scratch` (looked up on the first
// call on each device, with that device current).
static int noalloc_scratch(char** out) {
  int dev = 0;
  DH_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(DH_ERR_UNSUPPORTED, "dh_broadphase_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    DH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = p;
  }
  *out = static_cast<char*>(g_na_ptr[dev]);
  return DH_OK;
}

}  // namespace dh

using namespace dh;

extern "C" int dh_broadphase(const dh_geometry* geom, const int* obstacle_dev, double cell,
                             int* pairs_dev, int cap, int* npairs, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!geom || !npairs || !(cell > 0) || cap < 0 || (cap > 0 && !pairs_dev))
    return set_error(DH_ERR_ARG, "dh_broadphase: bad arguments");
  *npairs = 0;
  if (geom->nparticles < 2) return DH_OK;
  long long np = 0;
  rc = broadphase_impl(geom, obstacle_dev, cell, pairs_dev, cap, &np, as_stream(stream));
  *npairs = static_cast<int>(np);
  return rc;
}

extern "C" int dh_broadphase_noalloc(const dh_geometry* geom, const int* obstacle_dev,
                                     double cell, int* pairs_dev, int cap, int* npairs,
                                     void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!geom || !npairs || !(cell > 0) || cap < 0 || (cap > 0 && !pairs_dev) ||
      geom->nparticles < 0 || geom->ntri_total < 0)
    return set_error(DH_ERR_ARG, "dh_broadphase_noalloc: bad arguments");
  *npairs = 0;
  if (geom->nparticles > DH_NOALLOC_BP_PARTICLES || geom->ntri_total > DH_NOALLOC_BP_TRIANGLES)
    return set_error(DH_ERR_CAPACITY,
                     "dh_broadphase_noalloc: %d particles / %d triangles exceed the fixed "
                     "scratch (at most %d / %d)",
                     geom->nparticles, geom->ntri_total, DH_NOALLOC_BP_PARTICLES,
                     DH_NOALLOC_BP_TRIANGLES);
  if (geom->nparticles < 2) return DH_OK;
  char* base = nullptr;
  if ((rc = noalloc_scratch(&base))) return rc;
  BpBuffers b;
  if (noalloc_layout(base, geom->nparticles, geom->ntri_total, &b) > kNaScratchBytes)
    return set_error(DH_ERR_CAPACITY, "dh_broadphase_noalloc: scratch layout too large");
  long long np = 0;
  rc = broadphase_run(geom, obstacle_dev, cell, b, pairs_dev, cap, &np, as_stream(stream));
  *npairs = static_cast<int>(np);
  return rc;
}

// Bytes the fixed scratch must hold for the largest accepted sizes (a
// build-time check of kNaScratchBytes, exported for tests).
extern "C" size_t dh_broadphase_noalloc_scratch_bytes(int nparticles, int ntri) {
  return noalloc_layout(nullptr, nparticles, ntri, nullptr);
}
