// detect.hip — particle-pair narrow phase: the reference's drivers bf,
// penalty, hybridWithPer{Batch,Triangle}FallBack, penaltyStat and
// hybrid{Batch,Triangle}Stat (bf.cpp:28-149, penalty.cpp:50-241,
// hybrid.cpp:59-627) over a batch of particle pairs.
//
// A particle pair's rows (triangles of A) are split into `tiles` contiguous
// ranges, one workgroup (4 waves) per range, which walks its rows' triangle
// pairs in row-major chunks of 256.  Several workgroups per particle pair
// keep the CUs busy when there are few pairs (a 64-particle scene has a few
// hundred).  Row decisions of the hybrid drivers (how many penalty solves
// failed in row iA) are accumulated in LDS in a first pass; the emitting
// pass then evaluates each pair with the method its row requires.  Contacts
// are compacted per chunk with wave ballots and one global atomic per chunk,
// tagged with (tile, rank in tile), and scattered into pair order
// afterwards, so the output order equals the reference's serial emission
// order and does not depend on scheduling or on the tiling.
#include <cstring>
#include <vector>

#include "exact_math.hpp"
#include "gjk.hpp"
#include "runtime.hpp"

namespace dh {

// Longest row the reference's drivers support: per-row stack arrays of
// 10000 entries indexed by 3*iB (bf.cpp:98, penalty.cpp:186, hybrid.cpp:402).
constexpr int kMaxRowTriangles = 3333;

// LDS list of DH_BF candidates that pass the box test (per workgroup), and
// the most triangles per particle whose boxes are staged in LDS.
constexpr int kSurvivors = 1024;
constexpr int kBoxTriangles = kRowBlock;
constexpr int kChunkFlags = 1024;  // chunks of 64 rows an item tests up front

enum DetectError : int { DE_NONE = 0, DE_REF_UB = 1, DE_TOO_LONG = 2 };

// Per-tile counters (tile_stats) and the stats slot each one is summed into.
enum TileStat : int {
  TS_CHECKS = 0, TS_EVALUATED, TS_FALLBACK_ROWS, TS_PENALTY_FAILS, TS_BATCH_FAILS,
  TS_BATCH_SIZE, TS_CULLED, kTileStats
};

enum StatSlot : int {
  ST_HIST = 0,            // 17 slots
  ST_PENALTY_FAILS = 17,
  ST_BATCH_FAILS = 18,
  ST_BATCH_SIZE = 19,
  ST_CHECKS = 20,
  ST_EVALUATED = 21,
  ST_FALLBACK_ROWS = 22,
  ST_CULLED = 23,  // DH_BF triangle pairs skipped by the box test
  ST_COUNT = 24
};

struct DetectArgs {
  const double* x;
  const double* y;
  const double* z;
  const int* tri_offset;
  const double* eps;
  const int* friction;
  const int* id;
  const int* pairs;
  int npairs;
  dh_contact* temp;  // unordered contacts: pair = tile, pad_ = rank within the tile
  long long cap;
  unsigned long long* counter;  // [0] contacts reserved
  int tiles;                    // row ranges (workgroups) per particle pair
  int* tile_count;              // contacts per (pair, tile), pair-major
  // kTileStats counters per tile (TS_*), summed by k_tile_stats afterwards:
  // one global atomic per workgroup and counter would serialise the launch
  unsigned long long* tile_stats;
  unsigned long long* stats;  // ST_COUNT counters
  int* error;                 // DetectError of the first failing pair
  // When set, the number of pairs is read here (at most npairs, which is then
  // the launch capacity); pairs beyond it record no contacts.
  const long long* npairs_dev;
  // When set, a box per particle (lo xyz, hi xyz) containing its vertices:
  // the broad phase's inflated boxes.  Rows apart from B's particle box are
  // apart from every box of B's triangles, so a block of rows that is all
  // apart is skipped before B's triangles are loaded.
  const double* pbox;
  // With pbox, DH_BF: the particles' bounding spheres (x, y, z, r) of the
  // same broad phase.  A pair whose spheres are apart by more than the
  // margin (and a rounding band) has every triangle pair beyond the margin
  // and is skipped whole (culled), as the reference's engine loop may skip
  // it by its own bounding-sphere test (Engine.cpp:272-295).
  const double* psph;
  // With pbox: the broad phase's boxes of the 64-triangle chunks of the
  // particles with more than kRowBlock triangles (chunk c of particle i at
  // bbox[6 * (tri_offset[i] / 64 + i + c)]).  Rows of a chunk apart from
  // B's particle box are not loaded; a block without a near chunk is
  // skipped.
  const double* bbox;
  // Survivor queue (null: every item solves its own survivors).  A staged
  // single-block DH_BF item whose survivors all fit in its LDS list appends
  // them {item, iA | iB << 16} as one contiguous range and records its start
  // in qbase[item]; when the range does not fit in qcap entries, the item
  // fills the part inside with {-1, 0} and solves its survivors itself.
  int2* queue;
  long long qcap;
  unsigned long long* qcounter;
  long long* qbase;
  // Mid-phase-only launches (no solver in the kernel) append the items they
  // cannot queue here; a second launch runs those items in full, taking its
  // items from item_list.
  int* deferred;
  unsigned* ndeferred;
  // Items k_detect_wave takes for their rows' test but cannot stage (B has
  // more than kWaveRows triangles) and does not cull: k_detect_big runs them.
  int* biglist;
  unsigned* nbig;
  // Items k_lane_cull leaves for k_detect_wave (with wlist, k_detect_wave
  // takes its items from this list, *nwlist long).
  int* wlist;
  unsigned* nwlist;
  const int* item_list;
  const unsigned* nitems_dev;  // with item_list: its length (else the launch's)
  int fp32_screen;             // k_queue_solve screens in binary32 first
  int wave_items;              // small items go to k_detect_wave (mid-phase only)
  // item_skip[item] != 0: k_chunk_cull culled the item and wrote its
  // counters (a launch of DH_BF without the survivor queue skips it)
  const unsigned char* item_skip;
};

// Items of a mid-phase-only launch small enough for one wave each.
constexpr int kWaveRows = 80;       // rows and columns staged per wave
constexpr int kWaveSurvivors = 256; // survivors listed per wave
__device__ __forceinline__ bool wave_item(int nrows, int nA, int nB) {
  return nrows <= kWaveRows && nB <= kWaveRows && nA < 65536;
}
// Items whose rows k_detect_wave tests against B's particle box (a larger B,
// such as a wall against a small particle, is often culled there).
__device__ __forceinline__ bool wave_rows(int nrows, int nA, int nB) {
  return nrows <= kWaveRows && nA < 65536 && nB < 65536;
}

__device__ __forceinline__ Tri<double> load_tri(const DetectArgs& g, int t) {
  Tri<double> r;
  const int b = 3 * t;
#pragma unroll
  for (int k = 0; k < 3; ++k) r.v[k] = {g.x[b + k], g.y[b + k], g.z[b + k]};
  return r;
}

// Mid-phase of DH_BF: true when the triangles' bounding boxes, grown by the
// contact margin, are apart along some axis by more than a rounding band.
// The distance then exceeds the margin, so bf (bf.cpp:125-127) emits no
// contact and the pair need not be solved.  Coordinates that are NaN drop
// out of the boxes; bfSolver's result never comes from them either (its
// strict minimum skips NaN features), so the test stays exact.
__device__ __forceinline__ void tri_box(const Tri<double>& T, double* lo, double* hi) {
  lo[0] = fmin(fmin(T.v[0].x, T.v[1].x), T.v[2].x);
  lo[1] = fmin(fmin(T.v[0].y, T.v[1].y), T.v[2].y);
  lo[2] = fmin(fmin(T.v[0].z, T.v[1].z), T.v[2].z);
  hi[0] = fmax(fmax(T.v[0].x, T.v[1].x), T.v[2].x);
  hi[1] = fmax(fmax(T.v[0].y, T.v[1].y), T.v[2].y);
  hi[2] = fmax(fmax(T.v[0].z, T.v[1].z), T.v[2].z);
}

__device__ __forceinline__ bool boxes_apart(const double* loA, const double* hiA,
                                            const double* loB, const double* hiB, double margin) {
  double gap = -INFINITY, mag = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    gap = fmax(gap, fmax(loB[q] - hiA[q], loA[q] - hiB[q]));
    mag = fmax(mag, fmax(fmax(fabs(loA[q]), fabs(hiA[q])), fmax(fabs(loB[q]), fabs(hiB[q]))));
  }
  return gap > margin * (1.0 + 0x1p-20) + mag * 0x1p-40;
}

__device__ __forceinline__ bool boxes_apart(const Tri<double>& A, const Tri<double>& B,
                                            double margin) {
  double a[6], b[6];
  tri_box(A, a, a + 3);
  tri_box(B, b, b + 3);
  return boxes_apart(a, a + 3, b, b + 3, margin);
}

// Lanes below this one whose predicate is set (wave64).
__device__ __forceinline__ int lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0));
}

struct Pair {
  int pa, pb, oA, oB, nA, nB;
  double epsA, epsB, margin, max_err;
  int fric, idA, idB;
};

__device__ __forceinline__ Pair load_pair(const DetectArgs& g, int k) {
  Pair p;
  p.pa = g.pairs[2 * k];
  p.pb = g.pairs[2 * k + 1];
  p.oA = g.tri_offset[p.pa];
  p.oB = g.tri_offset[p.pb];
  p.nA = g.tri_offset[p.pa + 1] - p.oA;
  p.nB = g.tri_offset[p.pb + 1] - p.oB;
  p.epsA = g.eps[p.pa];
  p.epsB = g.eps[p.pb];
  p.margin = p.epsA + p.epsB;
  p.max_err = (p.epsA + p.epsB) / 16.0;
  p.fric = (g.friction[p.pa] && g.friction[p.pb]) ? 1 : 0;
  p.idA = g.id[p.pa];
  p.idB = g.id[p.pb];
  return p;
}

// The particles' bounding spheres (the broad phase's centre and largest
// vertex distance) apart by more than the margin, with an allowance for the
// rounding of the spheres and of bfSolver's distance: no triangle pair of
// the two particles emits a contact.
__device__ __forceinline__ bool spheres_apart(const DetectArgs& g, const Pair& p) {
  if (!g.psph) return false;
  const double* sa = g.psph + 4 * static_cast<long long>(p.pa);
  const double* sb = g.psph + 4 * static_cast<long long>(p.pb);
  const double dx = sa[0] - sb[0], dy = sa[1] - sb[1], dz = sa[2] - sb[2];
  const double mag = fmax(fmax(fmax(fabs(sa[0]), fabs(sa[1])), fmax(fabs(sa[2]), fabs(sb[0]))),
                          fmax(fabs(sb[1]), fabs(sb[2]))) + sa[3] + sb[3];
  const double reach = sa[3] + sb[3] + p.margin * (1.0 + 0x1p-20) + mag * 0x1p-36;
  return dx * dx + dy * dy + dz * dz > reach * reach;
}

// A triangle's box [lo, hi] apart from a particle's bounding sphere s
// (centre, radius) by more than the margin, with spheres_apart's allowance:
// no triangle of that particle is within the margin of the box's triangle.
__device__ __forceinline__ bool box_sphere_apart(const double* lo, const double* hi,
                                                 const double* s, double margin) {
  double d2 = 0.0, mag = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double e = fmax(fmax(lo[q] - s[q], s[q] - hi[q]), 0.0);
    d2 += e * e;
    mag = fmax(mag, fmax(fmax(fabs(lo[q]), fabs(hi[q])), fabs(s[q])));
  }
  const double reach = s[3] + margin * (1.0 + 0x1p-20) + (mag + s[3]) * 0x1p-36;
  return d2 > reach * reach;
}

__device__ __forceinline__ bool penalty_pair(const Tri<double>& A, const Tri<double>& B,
                                             double max_err, V3<double>& P, V3<double>& Q) {
  const double x0[4] = {0.33, 0.33, 0.33, 0.33};
  return penalty_fixed(A, B, x0, 4, max_err, P, Q, (double*)nullptr);
}

// Row fallback rules.
template <int M>
__device__ __forceinline__ bool row_falls_back(int fails, int nB, double max_err) {
  if (M == DH_HYBRID_TRIANGLE) return fails > 0;
  const double local = fails * max_err;
  if (M == DH_HYBRID_BATCH) return local > (max_err * static_cast<double>(3 * nB)) / 32.0;
  return local > (max_err * static_cast<double>(nB)) / 32;  // *Stat variants
}

template <int M>
__device__ __forceinline__ double final_pair(const DetectArgs& g, const Pair& p, int iA, int iB,
                                             const int* row_fail, int r0, V3<double>& P,
                                             V3<double>& Q) {
  const Tri<double> A = load_tri(g, p.oA + iA);
  const Tri<double> B = load_tri(g, p.oB + iB);
  if (M == DH_BF) {
    bf_solver(A, B, P, Q);
  } else if (M == DH_PENALTY) {
    penalty_pair(A, B, p.max_err, P, Q);
  } else if (M == DH_PENALTY_STAT) {
    penalty_adaptive(A, B, p.max_err, P, Q);
  } else {
    if (row_falls_back<M>(row_fail[iA - r0], p.nB, p.max_err))
      bf_solver(A, B, P, Q);
    else
      penalty_pair(A, B, p.max_err, P, Q);
  }
  return pq_dist(P, Q);
}

__device__ __forceinline__ void make_contact(const Pair& p, int k, int iA, int iB,
                                             const V3<double>& P, const V3<double>& Q,
                                             dh_contact* c) {
  // contactpoint.cpp:116-155
  c->x[0] = (P.x + Q.x) / 2.0;
  c->x[1] = (P.y + Q.y) / 2.0;
  c->x[2] = (P.z + Q.z) / 2.0;
  c->P[0] = P.x; c->P[1] = P.y; c->P[2] = P.z;
  c->Q[0] = Q.x; c->Q[1] = Q.y; c->Q[2] = Q.z;
  c->eps_total = p.epsA + p.epsB;
  const double dist = pq_dist(P, Q);
  c->depth = c->eps_total - dist;
  c->normal[0] = (P.x - Q.x) / dist;
  c->normal[1] = (P.y - Q.y) / dist;
  c->normal[2] = (P.z - Q.z) / dist;
  c->master = p.idA;
  c->slave = p.idB;
  c->penetration = 1;
  c->friction = p.fric;
  c->triA = iA;
  c->triB = iB;
  c->pair = k;
}

// One (particle pair, row range) item of the detection: `item` = pair *
// tiles + tile.  Returns when the item is done; the caller synchronises the
// workgroup before its next item.
template <int M, bool QONLY>
__device__ __forceinline__ void detect_item(const DetectArgs& g, const int item) {
  extern __shared__ unsigned char smem[];
  __shared__ int wave_cnt[4];
  __shared__ unsigned long long chunk_base;
  __shared__ unsigned long long local_stats[6];

  const int k = item / g.tiles, tile = item - k * g.tiles;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  unsigned long long* ts = g.tile_stats + static_cast<long long>(item) * kTileStats;
  if (M == DH_BF && !QONLY && g.item_skip && g.item_skip[item]) return;  // k_chunk_cull's
  if (g.npairs_dev && k >= *g.npairs_dev) {
    if (tid == 0) g.tile_count[item] = 0;
    if (tid < kTileStats) ts[tid] = 0;
    return;
  }
  const Pair p = load_pair(g, k);
  const bool is_hybrid = (M == DH_HYBRID_BATCH || M == DH_HYBRID_TRIANGLE ||
                          M == DH_HYBRID_BATCH_STAT || M == DH_HYBRID_TRIANGLE_STAT);
  const bool is_rowstat = (M == DH_HYBRID_BATCH_STAT || M == DH_HYBRID_TRIANGLE_STAT);
  const int rows = (M == DH_HYBRID_TRIANGLE) ? p.nB : p.nA;

  if (M == DH_HYBRID_TRIANGLE && p.nB > p.nA) {
    if (tid == 0) atomicCAS(g.error, DE_NONE, static_cast<int>(DE_REF_UB));
    if (tid == 0) g.tile_count[item] = 0;
    if (tid < kTileStats) ts[tid] = 0;
    return;
  }
  if ((is_hybrid || is_rowstat) && (rows > kMaxRowTriangles || p.nB > kMaxRowTriangles)) {
    if (tid == 0) atomicCAS(g.error, DE_NONE, static_cast<int>(DE_TOO_LONG));
    if (tid == 0) g.tile_count[item] = 0;
    if (tid < kTileStats) ts[tid] = 0;
    return;
  }
  // this workgroup's rows [r0, r1)
  const int r0 = static_cast<int>(static_cast<long long>(tile) * rows / g.tiles);
  const int r1 = static_cast<int>(static_cast<long long>(tile + 1) * rows / g.tiles);
  const int nrows = r1 - r0;
  if (tid < 6) local_stats[tid] = 0;

  int* row_fail = reinterpret_cast<int*>(smem);
  int* row_lastnan = row_fail + kMaxRowTriangles;
  unsigned long long* row_min =
      reinterpret_cast<unsigned long long*>(smem + 2 * kMaxRowTriangles * sizeof(int) + 8);
  const int total = nrows * p.nB;  // triangle pairs of this tile, from row r0
  const int f0 = r0 * p.nB;

  if (is_hybrid) {
    for (int r = tid; r < nrows; r += 256) {
      row_fail[r] = 0;
      if (is_rowstat) {
        row_lastnan[r] = -1;
        row_min[r] = ~0ull;
      }
    }
    __syncthreads();
    // pass 1: penalty on every pair, failures per row
    for (int f = f0 + tid; f < f0 + total; f += 256) {
      const int iA = f / p.nB, iB = f - iA * p.nB;
      const Tri<double> A = load_tri(g, p.oA + iA);
      const Tri<double> B = load_tri(g, p.oB + iB);
      V3<double> P, Q;
      if (penalty_pair(A, B, p.max_err, P, Q)) atomicAdd(&row_fail[iA - r0], 1);
    }
    __syncthreads();
    if (tid == 0) {
      unsigned long long fb = 0, fails = 0, bfails = 0;
      for (int r = 0; r < nrows; ++r) {
        fails += row_fail[r];
        if (row_falls_back<M>(row_fail[r], p.nB, p.max_err)) {
          fb++;
          bfails++;
        }
      }
      local_stats[0] = fb;
      local_stats[1] = fails;
      local_stats[2] = bfails;
    }
  }
  if (is_rowstat) {
    // pass 2: last NaN distance per row; pass 3: minimum after it
    // (std::min(d, minD) semantics of hybrid.cpp:316-318)
    for (int f = f0 + tid; f < f0 + total; f += 256) {
      const int iA = f / p.nB, iB = f - iA * p.nB;
      V3<double> P, Q;
      const double d = final_pair<M>(g, p, iA, iB, row_fail, r0, P, Q);
      if (d != d) atomicMax(&row_lastnan[iA - r0], iB);
    }
    __syncthreads();
    for (int f = f0 + tid; f < f0 + total; f += 256) {
      const int iA = f / p.nB, iB = f - iA * p.nB;
      if (iB <= row_lastnan[iA - r0]) continue;
      V3<double> P, Q;
      const double d = final_pair<M>(g, p, iA, iB, row_fail, r0, P, Q);
      atomicMin(&row_min[iA - r0], static_cast<unsigned long long>(__double_as_longlong(d)));
    }
    __syncthreads();
  }

  // Emission of one chunk of 256 candidates (in row-major order across
  // chunks): block prefix over the emitting lanes, one global atomic for the
  // chunk's slots, rank within the tile in pad_.
  unsigned long long running = 0;
  unsigned culled = 0;
  auto emit_chunk = [&](bool emit, int iA, int iB, const V3<double>& P, const V3<double>& Q) {
    const unsigned long long mask = __ballot(emit);
    const int pre = lane_prefix(mask);
    if (lane == 0) wave_cnt[wave] = __popcll(mask);
    __syncthreads();
    int off = 0, cnt = 0;
    for (int w = 0; w < 4; ++w) {
      off += (w < wave) ? wave_cnt[w] : 0;
      cnt += wave_cnt[w];
    }
    if (tid == 0 && cnt > 0) chunk_base = atomicAdd(g.counter, static_cast<unsigned long long>(cnt));
    __syncthreads();
    if (emit) {
      const unsigned long long slot = chunk_base + off + pre;
      if (slot < static_cast<unsigned long long>(g.cap)) {
        dh_contact c;
        make_contact(p, k, iA, iB, P, Q, &c);
        c.pair = item;  // the tile, until the scatter
        c.pad_ = static_cast<int>(running + off + pre);
        g.temp[slot] = c;
      }
    }
    running += cnt;
    __syncthreads();
  };

  if (M == DH_BF) {
    // Box test first, survivors compacted in row-major order into LDS, then
    // solved 256 at a time, so that no lane idles beside a culled pair.  When
    // B's triangle boxes fit in LDS they are staged there once, the tile's
    // rows are staged in blocks of up to kBoxTriangles (a large A such as a
    // funnel or a wall takes several blocks, in row order), and only rows
    // whose box meets the union of B's boxes and columns whose box meets the
    // union of the block's rows are paired (the union test implies the pair
    // test, so nothing else changes).  A larger B takes one unstaged pass.
    __shared__ int surv[kSurvivors];
    __shared__ double boxes[2 * kBoxTriangles * 6];
    __shared__ int act[2 * kBoxTriangles];  // active rows (block-relative), then columns
    __shared__ double uni[2][6];            // union of the block's row boxes, of B's boxes
    __shared__ double red[4][12];
    __shared__ int nact[2];
    __shared__ unsigned char row_far[kBoxTriangles];  // row in a chunk apart from B
    __shared__ unsigned char chunk_near[kChunkFlags];  // the item's chunks against B's box
    const bool staged = p.nB <= kBoxTriangles;  // block-uniform
    const int blk = staged ? kBoxTriangles : (nrows > 0 ? nrows : 1);
    bool b_staged = false;  // B's boxes are loaded on the first block that needs them
    long long q_first = 0;  // the item's queue range, when a block appended one
    int q_len = 0;
    // mid-phase-only launch: an item that cannot queue its survivors is
    // handed to the full launch (which also writes its counters)
    auto defer = [&]() {
      if (tid == 0) g.deferred[atomicAdd(g.ndeferred, 1u)] = item;
    };
    if (QONLY && g.wave_items && wave_item(nrows, p.nA, p.nB)) return;  // k_detect_wave's
    if (QONLY && !(staged && p.nA < 65536 && p.nB < 65536)) {
      defer();
      return;
    }
    // the whole pair beyond the margin by its bounding spheres
    const bool far_pair = spheres_apart(g, p);
    if (far_pair) culled += tid == 0 ? static_cast<unsigned>(nrows * p.nB) : 0u;
    // A in chunks of 64 rows with their boxes (a wall or a funnel): every
    // chunk of the item tested against B's particle box once, up front
    const bool chunked = staged && g.bbox && g.pbox && p.nA > kRowBlock;
    const double* pbB = g.pbox ? g.pbox + 6 * static_cast<long long>(p.pb) : nullptr;
    // row r of A (particle-relative) lies in a chunk apart from B's box
    auto chunk_far = [&](int r) {
      const double* cb = g.bbox + 6 * (static_cast<long long>(p.oA / 64) + p.pa + r / 64);
      return boxes_apart(cb, cb + 3, pbB, pbB + 3, p.margin);
    };
    const int c_first = r0 / 64;
    const bool flagged = chunked && nrows > 0 && (r0 + nrows - 1) / 64 - c_first < kChunkFlags;
    if (flagged && !far_pair) {
      for (int c = tid; c <= (r0 + nrows - 1) / 64 - c_first; c += 256)
        chunk_near[c] = chunk_far(64 * (c_first + c)) ? 0 : 1;
      __syncthreads();
    }
    auto row_chunk_far = [&](int r) {
      return flagged ? chunk_near[r / 64 - c_first] == 0 : chunk_far(r);
    };
    for (int rb = 0; !far_pair && rb < nrows; rb += blk) {
      const int rs = r0 + rb;  // first row of the block
      const int nrb = nrows - rb < blk ? nrows - rb : blk;
      int nra = nrb, nca = p.nB;
      if (flagged) {
        // the same flags in every thread: no barrier
        bool any_near = false;
        for (int c = rs / 64; c <= (rs + nrb - 1) / 64; ++c)
          any_near = any_near || chunk_near[c - c_first] != 0;
        if (!any_near) {  // every chunk of the block is apart
          culled += tid == 0 ? static_cast<unsigned>(nrb * p.nB) : 0u;
          continue;
        }
      } else if (chunked) {
        bool any_near = false;
        for (int c = rs / 64 + tid; c <= (rs + nrb - 1) / 64; c += 256)
          any_near = any_near || !chunk_far(64 * c);
        if (!__syncthreads_or(any_near)) {  // every chunk of the block is apart
          culled += tid == 0 ? static_cast<unsigned>(nrb * p.nB) : 0u;
          continue;
        }
      }
      if (staged) {
        bool near = false;
        for (int t = tid; t < nrb; t += 256) {
          const bool far = chunked && row_chunk_far(rs + t);
          row_far[t] = far ? 1 : 0;
          if (far) continue;  // its box is never read
          double* b = boxes + 6 * t;
          tri_box(load_tri(g, p.oA + rs + t), b, b + 3);
          if (g.pbox) near = near || !boxes_apart(b, b + 3, pbB, pbB + 3, p.margin);
        }
        if (g.pbox && !__syncthreads_or(near)) {
          // every row of the block misses B's particle box
          culled += tid == 0 ? static_cast<unsigned>(nrb * p.nB) : 0u;
          continue;
        }
        if (!b_staged) {
          for (int t = tid; t < p.nB; t += 256) {
            double* b = boxes + 6 * (kBoxTriangles + t);
            tri_box(load_tri(g, p.oB + t), b, b + 3);
          }
          b_staged = true;
        }
        __syncthreads();
        // unions: lanes over strided boxes, then waves, then the block
        double u[12];
        for (int q = 0; q < 3; ++q) {
          u[q] = u[6 + q] = INFINITY;
          u[3 + q] = u[9 + q] = -INFINITY;
        }
        for (int t = tid; t < nrb; t += 256)
          if (!row_far[t])
            for (int q = 0; q < 3; ++q) {
              u[q] = fmin(u[q], boxes[6 * t + q]);
              u[3 + q] = fmax(u[3 + q], boxes[6 * t + 3 + q]);
            }
        for (int t = tid; t < p.nB; t += 256)
          for (int q = 0; q < 3; ++q) {
            u[6 + q] = fmin(u[6 + q], boxes[6 * (kBoxTriangles + t) + q]);
            u[9 + q] = fmax(u[9 + q], boxes[6 * (kBoxTriangles + t) + 3 + q]);
          }
        for (int o = 32; o > 0; o >>= 1)
          for (int q = 0; q < 12; ++q) {
            const double v = __shfl_xor(u[q], o);
            u[q] = ((q % 6) < 3) ? fmin(u[q], v) : fmax(u[q], v);
          }
        if (lane == 0)
          for (int q = 0; q < 12; ++q) red[wave][q] = u[q];
        __syncthreads();
        if (tid < 12) {
          double v = red[0][tid];
          for (int w = 1; w < 4; ++w) v = ((tid % 6) < 3) ? fmin(v, red[w][tid]) : fmax(v, red[w][tid]);
          uni[tid / 6][tid % 6] = v;
        }
        if (tid < 2) nact[tid] = 0;
        __syncthreads();
        // ordered lists of active rows and columns
        for (int side = 0; side < 2; ++side) {
          const int cnt_all = side == 0 ? nrb : p.nB;
          const double* other = uni[1 - side];
          for (int t0 = 0; t0 < cnt_all; t0 += 256) {
            const int t = t0 + tid;
            bool on = false;
            if (t < cnt_all && !(side == 0 && row_far[t])) {
              const double* b = boxes + 6 * (side == 0 ? t : kBoxTriangles + t);
              on = !boxes_apart(b, b + 3, other, other + 3, p.margin);
            }
            const unsigned long long m = __ballot(on);
            if (lane == 0) wave_cnt[wave] = __popcll(m);
            __syncthreads();
            int off = 0, cnt = 0;
            for (int w = 0; w < 4; ++w) {
              off += (w < wave) ? wave_cnt[w] : 0;
              cnt += wave_cnt[w];
            }
            if (on) act[side * kBoxTriangles + nact[side] + off + lane_prefix(m)] = t;
            __syncthreads();
            if (tid == 0) nact[side] += cnt;
            __syncthreads();
          }
        }
        nra = nact[0];
        nca = nact[1];
      }
      // candidate fe in [0, nra * nca): row-major over the active rows and columns
      auto pair_of = [&](int fe, int& iA, int& iB) {
        const int ra = fe / nca, cb = fe - ra * nca;
        iA = staged ? rs + act[ra] : rs + ra;
        iB = staged ? act[kBoxTriangles + cb] : cb;
      };
      const int tested = nra * nca;
      culled += tid == 0 ? static_cast<unsigned>(nrb * p.nB - tested) : 0u;
      for (int fpos = 0; fpos < tested;) {
        int nl = 0;
        const bool first = fpos == 0;
        while (fpos < tested && nl <= kSurvivors - 256) {
          const int fe = fpos + tid;
          bool keep = false;
          if (fe < tested) {
            int iA, iB;
            pair_of(fe, iA, iB);
            if (staged) {
              const double* a = boxes + 6 * (iA - rs);
              const double* b = boxes + 6 * (kBoxTriangles + iB);
              keep = !boxes_apart(a, a + 3, b, b + 3, p.margin);
            } else {
              keep = !boxes_apart(load_tri(g, p.oA + iA), load_tri(g, p.oB + iB), p.margin);
            }
            culled += keep ? 0u : 1u;
          }
          const unsigned long long m = __ballot(keep);
          if (lane == 0) wave_cnt[wave] = __popcll(m);
          __syncthreads();
          int off = 0, cnt = 0;
          for (int w = 0; w < 4; ++w) {
            off += (w < wave) ? wave_cnt[w] : 0;
            cnt += wave_cnt[w];
          }
          if (keep) surv[nl + off + lane_prefix(m)] = fe;
          nl += cnt;
          fpos += 256;
          __syncthreads();
        }
        if (QONLY && fpos < tested) {  // more survivors than the LDS list holds
          defer();
          return;
        }
        if (g.queue && first && fpos >= tested && staged && p.nA < 65536 && p.nB < 65536 &&
            !(nl > 0 && q_len > 0)) {
          // every survivor of the block is in LDS: append them, in row-major
          // order, as the item's one range of the survivor queue (solved
          // there with every lane busy); the item solves them itself if they
          // do not fit
          if (tid == 0 && nl > 0)
            chunk_base = atomicAdd(g.qcounter, static_cast<unsigned long long>(nl));
          __syncthreads();
          const long long qb = nl > 0 ? static_cast<long long>(chunk_base) : 0;
          __syncthreads();
          if (qb + nl <= g.qcap) {
            if (nl > 0) {
              q_first = qb;
              q_len = nl;
            }
            if (tid == 0 && nl > 0) g.qbase[item] = qb;
            for (int e = tid; e < nl; e += 256) {
              int iA, iB;
              pair_of(surv[e], iA, iB);
              g.queue[qb + e] = make_int2(
                  item, static_cast<int>(static_cast<unsigned>(iA) | (static_cast<unsigned>(iB) << 16)));
            }
            continue;  // fpos >= tested: the item is done
          }
          for (long long q = qb + tid; q < qb + nl && q < g.qcap; q += 256)
            g.queue[q] = make_int2(-1, 0);
          if (QONLY) {
            defer();
            return;
          }
        }
        if (QONLY && nl > 0 && q_len > 0) {
          // a second block with survivors: the item's survivors would not
          // be one range; void the first and run the item in full
          for (long long q = q_first + tid; q < q_first + q_len; q += 256)
            g.queue[q] = make_int2(-1, 0);
          defer();
          return;
        }
        if (!QONLY) {
          // the survivors that a triangle plane separates by more than the
          // margin (and, with the binary32 screen, those it proves far) drop
          // out; the rest are compacted in place (order kept) for the
          // binary64 solve.  (The edge-pair axes of the queue paths do not
          // pay here: an item keeps one wave's bfSolver round as long as a
          // survivor is left, whatever their number.)
          int n2 = 0;
          for (int b = 0; b < nl; b += 256) {
            const int e = b + tid;
            int v = 0;
            bool keep2 = false;
            if (e < nl) {
              v = surv[e];
              int iA, iB;
              pair_of(v, iA, iB);
              const Tri<double> A = load_tri(g, p.oA + iA), B = load_tri(g, p.oB + iB);
              keep2 = !plane_apart(A, B, p.margin) &&
                      !(g.fp32_screen && screened_far(A, B, p.margin));
            }
            const unsigned long long m = __ballot(keep2);
            if (lane == 0) wave_cnt[wave] = __popcll(m);
            __syncthreads();
            int off = 0, cnt = 0;
            for (int w = 0; w < 4; ++w) {
              off += (w < wave) ? wave_cnt[w] : 0;
              cnt += wave_cnt[w];
            }
            if (keep2) surv[n2 + off + lane_prefix(m)] = v;
            n2 += cnt;
            __syncthreads();
          }
          nl = n2;
        }
        for (int b = 0; b < (QONLY ? 0 : nl); b += 256) {
          const int e = b + tid;
          bool emit = false;
          V3<double> P = {0, 0, 0}, Q = {0, 0, 0};
          int iA = 0, iB = 0;
          if (e < nl) {
            pair_of(surv[e], iA, iB);
            bf_solver(load_tri(g, p.oA + iA), load_tri(g, p.oB + iB), P, Q);
            emit = pq_dist(P, Q) <= p.margin;
          }
          emit_chunk(emit, iA, iB, P, Q);
        }
      }
      // the next block overwrites the row boxes and the active lists
      __syncthreads();
    }
  } else {
    for (int base = 0; base < total; base += 256) {
      const int f = base + tid;
      bool emit = false;
      V3<double> P = {0, 0, 0}, Q = {0, 0, 0};
      int iA = 0, iB = 0;
      if (f < total) {
        iA = (f0 + f) / p.nB;
        iB = (f0 + f) - iA * p.nB;
        if (M == DH_PENALTY_STAT) {
          const Tri<double> A = load_tri(g, p.oA + iA);
          const Tri<double> B = load_tri(g, p.oB + iB);
          const int its = penalty_adaptive(A, B, p.max_err, P, Q);
          atomicAdd(&g.stats[ST_HIST + (its < 0 ? 0 : its)], 1ull);
          const double d = pq_dist(P, Q);
          emit = (3 * iB < p.nB) && (d <= p.margin);
        } else {
          const double d = final_pair<M>(g, p, iA, iB, row_fail, r0, P, Q);
          if (is_rowstat) {
            double mn;
            if (row_lastnan[iA - r0] == p.nB - 1) {
              mn = __longlong_as_double(0x7ff8000000000000ll);
            } else {
              mn = __longlong_as_double(static_cast<long long>(row_min[iA - r0]));
              if (row_lastnan[iA - r0] < 0) mn = (1E99 < mn) ? 1E99 : mn;
            }
            emit = (d <= mn) && (d < p.margin);
          } else {
            emit = d < p.margin;
          }
        }
      }
      emit_chunk(emit, iA, iB, P, Q);
    }
  }
  if (M == DH_BF) {
    // slot 3 of local_stats (zeroed at the start, unused by BF otherwise)
    for (int o = 32; o > 0; o >>= 1) culled += __shfl_xor(culled, o);
    if (lane == 0) atomicAdd(&local_stats[3], static_cast<unsigned long long>(culled));
    __syncthreads();
  }
  if (tid == 0) {
    g.tile_count[item] = static_cast<int>(running);
    // evaluations the reference performs: every pair once, plus bfSolver
    // again over each fallback row (sums over the tiles)
    const unsigned long long evals =
        static_cast<unsigned long long>(total) +
        (is_hybrid ? local_stats[0] * static_cast<unsigned long long>(p.nB) : 0ull);
    ts[TS_CHECKS] = tile == 0 ? static_cast<unsigned long long>(p.nA) * p.nB : 0ull;
    ts[TS_EVALUATED] = evals;
    ts[TS_FALLBACK_ROWS] = is_hybrid ? local_stats[0] : 0ull;
    ts[TS_PENALTY_FAILS] = is_rowstat ? local_stats[1] : 0ull;
    ts[TS_BATCH_FAILS] = is_rowstat ? local_stats[2] : 0ull;
    ts[TS_BATCH_SIZE] = is_rowstat ? static_cast<unsigned long long>(nrows) * p.nB : 0ull;
    ts[TS_CULLED] = M == DH_BF ? local_stats[3] : 0ull;
  }
}

// Persistent over the items: workgroup b takes items b, b + gridDim.x, ...
// (with as many workgroups as items, one each).  Pairs of very uneven cost
// (a 2,048-triangle funnel against 20-triangle particles) share the
// workgroups of one launch.  The brute force is held to 168 registers
// (three waves per SIMD; a few spill), which its latency-bound item loop
// needs more than the registers.
template <int M, bool QONLY = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(M == DH_BF && !QONLY ? 3 : 1))) void k_detect(DetectArgs g,
                                                                                 int items) {
  if (g.item_list && g.nitems_dev && static_cast<int>(*g.nitems_dev) < items)
    items = static_cast<int>(*g.nitems_dev);
  for (int t = blockIdx.x; t < items; t += gridDim.x) {
    detect_item<M, QONLY>(g, g.item_list ? g.item_list[t] : t);
    __syncthreads();  // the next item reuses the workgroup's LDS
  }
}

// The workgroup mid-phase for the items k_detect_wave leaves (larger ones):
// the first wave classifies 64 items at a time, one per lane, and the
// workgroup runs the larger live ones, in order.
__global__ __launch_bounds__(256) void k_detect_big(DetectArgs g, int items) {
  __shared__ int list[64];
  __shared__ int nlist;
  const int tid = threadIdx.x;
  for (int base = blockIdx.x * 64; base < items; base += gridDim.x * 64) {
    if (tid < 64) {
      const int item = base + tid;
      bool big = false;
      if (item < items) {
        const int k = item / g.tiles, tile = item - k * g.tiles;
        if (!(g.npairs_dev && k >= *g.npairs_dev)) {
          const int pa = g.pairs[2 * k], pb = g.pairs[2 * k + 1];
          const int nA = g.tri_offset[pa + 1] - g.tri_offset[pa];
          const int nB = g.tri_offset[pb + 1] - g.tri_offset[pb];
          const int r0 = static_cast<int>(static_cast<long long>(tile) * nA / g.tiles);
          const int r1 = static_cast<int>(static_cast<long long>(tile + 1) * nA / g.tiles);
          big = !wave_rows(r1 - r0, nA, nB);
        }
      }
      const unsigned long long m = __ballot(big);
      if (big) list[lane_prefix(m)] = item;
      if (tid == 0) nlist = __popcll(m);
    }
    __syncthreads();
    const int n = nlist;
    for (int j = 0; j < n; ++j) {
      detect_item<DH_BF, true>(g, list[j]);
      __syncthreads();  // the next item reuses the workgroup's LDS
    }
  }
  // then the items k_detect_wave (launched before) handed over
  const int nb = static_cast<int>(*g.nbig);
  for (int j = blockIdx.x; j < nb; j += gridDim.x) {
    detect_item<DH_BF, true>(g, g.biglist[j]);
    __syncthreads();
  }
}

// One wave per item, before a DH_BF launch without the survivor queue, on
// scenes with particles of more than kRowBlock triangles (walls, funnels):
// an item whose A is cut in chunks of 64 rows is culled when no row of a
// chunk near B's particle box is itself near it, with the counters
// detect_item would write (its blocks would all be skipped).  The test is
// detect_item's chunk and row tests, without its workgroup barriers.
// item_skip[item] = 1 marks the culled and dead items.
__global__ __launch_bounds__(256) void k_chunk_cull(DetectArgs g, int items,
                                                    unsigned char* __restrict__ skip) {
  const int lane = threadIdx.x & 63;
  const int nw = static_cast<int>(gridDim.x) * 4;
  for (int item = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
       item < items; item += nw) {
    const int k = item / g.tiles, tile = item - k * g.tiles;
    unsigned long long* ts = g.tile_stats + static_cast<long long>(item) * kTileStats;
    if (g.npairs_dev && k >= *g.npairs_dev) {
      if (lane == 0) {
        g.tile_count[item] = 0;
        skip[item] = 1;
      }
      if (lane < kTileStats) ts[lane] = 0;
      continue;
    }
    const Pair p = load_pair(g, k);
    const int r0 = static_cast<int>(static_cast<long long>(tile) * p.nA / g.tiles);
    const int r1 = static_cast<int>(static_cast<long long>(tile + 1) * p.nA / g.tiles);
    const bool chunked = p.nB <= kBoxTriangles && g.bbox && g.pbox && p.nA > kRowBlock;
    bool near = !(chunked && r1 > r0);
    if (!near) {
      const double* pbB = g.pbox + 6 * static_cast<long long>(p.pb);
      const int c_last = (r1 - 1) / 64;
      for (int c0 = r0 / 64; c0 <= c_last && !near; c0 += 64) {
        const int c = c0 + lane;
        bool cn = false;
        if (c <= c_last) {
          const double* cb = g.bbox + 6 * (static_cast<long long>(p.oA / 64) + p.pa + c);
          cn = !boxes_apart(cb, cb + 3, pbB, pbB + 3, p.margin);
        }
        unsigned long long m = __ballot(cn);
        while (m && !near) {
          const int cc = c0 + __builtin_ctzll(m);
          m &= m - 1;
          const int r = 64 * cc + lane;
          bool rn = false;
          if (r >= r0 && r < r1) {
            double lo[3], hi[3];
            tri_box(load_tri(g, p.oA + r), lo, hi);
            rn = !boxes_apart(lo, hi, pbB, pbB + 3, p.margin);
          }
          near = __ballot(rn) != 0;
        }
      }
    }
    if (lane == 0) {
      skip[item] = near ? 0 : 1;
      if (!near) {
        const unsigned long long total = static_cast<unsigned long long>(r1 - r0) * p.nB;
        g.tile_count[item] = 0;
        ts[TS_CHECKS] = tile == 0 ? static_cast<unsigned long long>(p.nA) * p.nB : 0ull;
        ts[TS_EVALUATED] = total;
        ts[TS_FALLBACK_ROWS] = 0;
        ts[TS_PENALTY_FAILS] = 0;
        ts[TS_BATCH_FAILS] = 0;
        ts[TS_BATCH_SIZE] = 0;
        ts[TS_CULLED] = total;
      }
    }
  }
}

// Rows a lane tests alone in k_lane_cull.
constexpr int kLaneRows = 16;

// One lane per item, before k_detect_wave: an item of k_detect_wave's whose
// particles' bounding spheres are apart by more than the margin (diagonal
// neighbours in a packing), or with at most kLaneRows rows whose rows' boxes
// all miss B's particle box (grown by the margin: a wall of few triangles
// against the many particles inside it), is culled here, with the counters
// k_detect_wave would write; the other items of k_detect_wave's are listed
// for it.
__global__ __launch_bounds__(256) void k_lane_cull(DetectArgs g, int items) {
  const int item = blockIdx.x * 256 + static_cast<int>(threadIdx.x);
  bool listed = false;
  if (item < items) {
    const int k = item / g.tiles, tile = item - k * g.tiles;
    unsigned long long* ts = g.tile_stats + static_cast<long long>(item) * kTileStats;
    if (g.npairs_dev && k >= *g.npairs_dev) {
      g.tile_count[item] = 0;
      for (int q = 0; q < kTileStats; ++q) ts[q] = 0;
    } else {
      const Pair p = load_pair(g, k);
      const int r0 = static_cast<int>(static_cast<long long>(tile) * p.nA / g.tiles);
      const int r1 = static_cast<int>(static_cast<long long>(tile + 1) * p.nA / g.tiles);
      const int nrows = r1 - r0;
      if (wave_rows(nrows, p.nA, p.nB)) {
        bool near = !(g.pbox && nrows <= kLaneRows);
        if (spheres_apart(g, p)) {
          near = false;  // culled whole
        } else if (!near) {
          const double* pb = g.pbox + 6 * static_cast<long long>(p.pb);
          for (int t = 0; t < nrows && !near; ++t) {
            double lo[3], hi[3];
            tri_box(load_tri(g, p.oA + r0 + t), lo, hi);
            near = !boxes_apart(lo, hi, pb, pb + 3, p.margin);
          }
        }
        if (near) {
          listed = true;
        } else {
          const unsigned total = static_cast<unsigned>(nrows * p.nB);
          g.tile_count[item] = 0;
          ts[TS_CHECKS] = tile == 0 ? static_cast<unsigned long long>(p.nA) * p.nB : 0ull;
          ts[TS_EVALUATED] = total;
          ts[TS_FALLBACK_ROWS] = 0;
          ts[TS_PENALTY_FAILS] = 0;
          ts[TS_BATCH_FAILS] = 0;
          ts[TS_BATCH_SIZE] = 0;
          ts[TS_CULLED] = total;
        }
      }
    }
  }
  const unsigned long long m = __ballot(listed);
  if (m) {
    unsigned base = 0;
    if ((threadIdx.x & 63) == static_cast<unsigned>(__builtin_ctzll(m)))
      base = atomicAdd(g.nwlist, static_cast<unsigned>(__popcll(m)));
    base = __shfl(base, __builtin_ctzll(m));
    if (listed) g.wlist[base + lane_prefix(m)] = item;
  }
}

// The mid-phase of DH_BF for small items, one wave per item and no
// workgroup barrier: A's rows and B's columns (at most kWaveRows each)
// staged as boxes in the wave's LDS, rows apart from B's particle box
// dropped (an item with a larger B that this test does not cull goes to
// k_detect_big), rows and columns prefiltered against the other side's union,
// the surviving pairs listed in row-major order and appended to the
// survivor queue as one range.  Items whose survivors overflow the list or
// the queue are deferred to the full kernel.  Same tests, order and
// counters as detect_item's queue path.
__global__ __launch_bounds__(128) void k_detect_wave(DetectArgs g, int items) {
  constexpr int kW = 2;  // waves per workgroup
  __shared__ double boxes[kW][2 * kWaveRows * 6];
  __shared__ unsigned char act[kW][2 * kWaveRows];
  __shared__ int surv[kW][kWaveSurvivors];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double* bx = boxes[w];
  const int nw = g.wlist ? static_cast<int>(*g.nwlist) : items;
  for (int j = blockIdx.x * kW + w; j < nw; j += gridDim.x * kW) {
    const int item = g.wlist ? g.wlist[j] : j;
    const int k = item / g.tiles, tile = item - k * g.tiles;
    unsigned long long* ts = g.tile_stats + static_cast<long long>(item) * kTileStats;
    if (g.npairs_dev && k >= *g.npairs_dev) {
      if (lane == 0) g.tile_count[item] = 0;
      if (lane < kTileStats) ts[lane] = 0;
      continue;
    }
    const Pair p = load_pair(g, k);
    const int r0 = static_cast<int>(static_cast<long long>(tile) * p.nA / g.tiles);
    const int r1 = static_cast<int>(static_cast<long long>(tile + 1) * p.nA / g.tiles);
    const int nrows = r1 - r0;
    if (!wave_rows(nrows, p.nA, p.nB)) continue;  // the workgroup kernel's
    const int total = nrows * p.nB;
    unsigned culled = 0;  // lane 0's count is the item's
    const double* pb = g.pbox ? g.pbox + 6 * static_cast<long long>(p.pb) : nullptr;
    // rows: boxes, and the test against B's particle box
    bool near = false;
    for (int t = lane; t < nrows; t += 64) {
      double* b = bx + 6 * t;
      tri_box(load_tri(g, p.oA + r0 + t), b, b + 3);
      if (pb) near = near || !boxes_apart(b, b + 3, pb, pb + 3, p.margin);
    }
    int S = 0;
    bool deferred = false;
    if (pb && !__ballot(near)) {
      culled = total;
    } else if (p.nB > kWaveRows) {
      if (lane == 0) g.biglist[atomicAdd(g.nbig, 1u)] = item;
      continue;
    } else {
      for (int t = lane; t < p.nB; t += 64) {
        double* b = bx + 6 * (kWaveRows + t);
        tri_box(load_tri(g, p.oB + t), b, b + 3);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // unions of the rows' and of B's boxes
      double u[12];
      for (int q = 0; q < 3; ++q) {
        u[q] = u[6 + q] = INFINITY;
        u[3 + q] = u[9 + q] = -INFINITY;
      }
      for (int t = lane; t < nrows; t += 64)
        for (int q = 0; q < 3; ++q) {
          u[q] = fmin(u[q], bx[6 * t + q]);
          u[3 + q] = fmax(u[3 + q], bx[6 * t + 3 + q]);
        }
      for (int t = lane; t < p.nB; t += 64)
        for (int q = 0; q < 3; ++q) {
          u[6 + q] = fmin(u[6 + q], bx[6 * (kWaveRows + t) + q]);
          u[9 + q] = fmax(u[9 + q], bx[6 * (kWaveRows + t) + 3 + q]);
        }
      for (int o = 32; o > 0; o >>= 1)
        for (int q = 0; q < 12; ++q) {
          const double v = __shfl_xor(u[q], o);
          u[q] = ((q % 6) < 3) ? fmin(u[q], v) : fmax(u[q], v);
        }
      // ordered active rows and columns: boxes that meet the other side's
      // union and are not apart from its particle's bounding sphere
      int na = 0, nc = 0;
      for (int side = 0; side < 2; ++side) {
        const int cnt_all = side == 0 ? nrows : p.nB;
        const double* other = u + 6 * (1 - side);
        const double* osph =
            g.psph ? g.psph + 4 * static_cast<long long>(side == 0 ? p.pb : p.pa) : nullptr;
        int n_on = 0;
        for (int t0 = 0; t0 < cnt_all; t0 += 64) {
          const int t = t0 + lane;
          bool on = false;
          if (t < cnt_all) {
            const double* b = bx + 6 * (side == 0 ? t : kWaveRows + t);
            on = !boxes_apart(b, b + 3, other, other + 3, p.margin) &&
                 !(osph && box_sphere_apart(b, b + 3, osph, p.margin));
          }
          const unsigned long long m = __ballot(on);
          if (on) act[w][side * kWaveRows + n_on + lane_prefix(m)] = static_cast<unsigned char>(t);
          n_on += __popcll(m);
        }
        if (side == 0) na = n_on; else nc = n_on;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int tested = na * nc;
      culled = static_cast<unsigned>(total - tested);
      // survivors in row-major order over the active rows and columns
      for (int f0 = 0; f0 < tested; f0 += 64) {
        const int fe = f0 + lane;
        bool keep = false;
        if (fe < tested) {
          const int ra = fe / nc, cb = fe - ra * nc;
          const double* a = bx + 6 * act[w][ra];
          const double* b = bx + 6 * (kWaveRows + act[w][kWaveRows + cb]);
          keep = !boxes_apart(a, a + 3, b, b + 3, p.margin);
        }
        const unsigned long long m = __ballot(keep);
        if (keep && S + lane_prefix(m) < kWaveSurvivors) surv[w][S + lane_prefix(m)] = fe;
        S += __popcll(m);
        culled += static_cast<unsigned>(__popcll(__ballot(fe < tested && !keep)));
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (S > 0 && S <= kWaveSurvivors) {
        // survivors that a triangle plane or an edge-pair axis separates by
        // more than the margin emit nothing (k_queue_solve's first test):
        // only the rest are queued, compacted in place in row-major order
        int S2 = 0;
        for (int e0 = 0; e0 < S; e0 += 64) {
          const int e = e0 + lane;
          int fe = 0;
          bool keep = false;
          if (e < S) {
            fe = surv[w][e];
            const int ra = fe / nc, cb = fe - ra * nc;
            keep = !axes_apart(load_tri(g, p.oA + r0 + act[w][ra]),
                               load_tri(g, p.oB + act[w][kWaveRows + cb]), p.margin);
          }
          const unsigned long long m = __ballot(keep);
          if (keep) surv[w][S2 + lane_prefix(m)] = fe;
          S2 += __popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        S = S2;
      }
      if (S > kWaveSurvivors) {
        deferred = true;
      } else if (S > 0) {
        unsigned long long qb0 = 0;
        if (lane == 0) qb0 = atomicAdd(g.qcounter, static_cast<unsigned long long>(S));
        const long long qb = static_cast<long long>(__shfl(qb0, 0));
        if (qb + S <= g.qcap) {
          if (lane == 0) g.qbase[item] = qb;
          for (int e = lane; e < S; e += 64) {
            const int fe = surv[w][e];
            const int ra = fe / nc, cb = fe - ra * nc;
            const int iA = r0 + act[w][ra], iB = act[w][kWaveRows + cb];
            g.queue[qb + e] = make_int2(
                item, static_cast<int>(static_cast<unsigned>(iA) | (static_cast<unsigned>(iB) << 16)));
          }
        } else {
          for (long long q = qb + lane; q < qb + S && q < g.qcap; q += 64)
            g.queue[q] = make_int2(-1, 0);
          deferred = true;
        }
      }
      // the next item overwrites the wave's LDS
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (deferred) {
      if (lane == 0) g.deferred[atomicAdd(g.ndeferred, 1u)] = item;
      continue;
    }
    if (lane == 0) {
      g.tile_count[item] = 0;
      ts[TS_CHECKS] = tile == 0 ? static_cast<unsigned long long>(p.nA) * p.nB : 0ull;
      ts[TS_EVALUATED] = static_cast<unsigned long long>(total);
      ts[TS_FALLBACK_ROWS] = 0;
      ts[TS_PENALTY_FAILS] = 0;
      ts[TS_BATCH_FAILS] = 0;
      ts[TS_BATCH_SIZE] = 0;
      ts[TS_CULLED] = culled;
    }
  }
}

// Sums of the per-tile counters into the stats slots (one atomic per
// workgroup and counter).
__global__ __launch_bounds__(256) void k_tile_stats(const unsigned long long* __restrict__ ts,
                                                    long long ntiles,
                                                    unsigned long long* __restrict__ stats) {
  __shared__ unsigned long long red[kTileStats];
  if (threadIdx.x < kTileStats) red[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long acc[kTileStats] = {};
  for (long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; t < ntiles;
       t += static_cast<long long>(gridDim.x) * 256)
    for (int q = 0; q < kTileStats; ++q) acc[q] += ts[t * kTileStats + q];
  for (int q = 0; q < kTileStats; ++q) {
    unsigned long long v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&red[q], v);
  }
  __syncthreads();
  if (threadIdx.x < kTileStats && red[threadIdx.x]) {
    const int slot[kTileStats] = {ST_CHECKS, ST_EVALUATED, ST_FALLBACK_ROWS, ST_PENALTY_FAILS,
                                  ST_BATCH_FAILS, ST_BATCH_SIZE, ST_CULLED};
    atomicAdd(&stats[slot[threadIdx.x]], red[threadIdx.x]);
  }
}

// The tail of a captured detection over at most kTailPairs pairs, in one
// workgroup in place of three launches: the tile counters' sums (as
// k_tile_stats), the per-pair contact counts (as k_tile_sums; with one tile
// per pair the detection wrote them directly) and their exclusive prefix
// sums with the total at offs[npairs] (as scan_counts).
constexpr int kTailPairs = 4096;

__global__ __launch_bounds__(1024) void k_detect_tail_small(
    const unsigned long long* __restrict__ ts, long long ntiles,
    unsigned long long* __restrict__ stats, const int* __restrict__ tile_count, int tiles,
    int npairs, int* __restrict__ counts, long long* __restrict__ offs) {
  __shared__ unsigned long long red[kTileStats];
  __shared__ long long s_wave[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t < kTileStats) red[t] = 0;
  __syncthreads();
  unsigned long long acc[kTileStats] = {};
  for (long long k = t; k < ntiles; k += 1024)
    for (int q = 0; q < kTileStats; ++q) acc[q] += ts[k * kTileStats + q];
  for (int q = 0; q < kTileStats; ++q) {
    unsigned long long v = acc[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0 && v) atomicAdd(&red[q], v);
  }
  // four consecutive pairs per thread
  int c[4];
  long long local = 0;
  for (int u = 0; u < 4; ++u) {
    const int k = 4 * t + u;
    int v = 0;
    if (k < npairs) {
      if (tiles > 1) {
        for (int q = 0; q < tiles; ++q) v += tile_count[static_cast<long long>(k) * tiles + q];
        counts[k] = v;
      } else {
        v = counts[k];
      }
    }
    c[u] = v;
    local += v;
  }
  long long incl = local;
  for (int o = 1; o < 64; o <<= 1) {
    const long long u = __shfl_up(incl, o);
    incl += lane >= o ? u : 0;
  }
  if (lane == 63) s_wave[w] = incl;
  __syncthreads();
  long long before = 0;
  for (int q = 0; q < w; ++q) before += s_wave[q];
  long long run = before + incl - local;
  for (int u = 0; u < 4; ++u) {
    const int k = 4 * t + u;
    if (k < npairs) offs[k] = run;
    run += c[u];
  }
  if (t == 1023) offs[npairs] = before + incl;
  if (t < kTileStats && red[t]) {
    const int slot[kTileStats] = {ST_CHECKS, ST_EVALUATED, ST_FALLBACK_ROWS, ST_PENALTY_FAILS,
                                  ST_BATCH_FAILS, ST_BATCH_SIZE, ST_CULLED};
    atomicAdd(&stats[slot[t]], red[t]);
  }
}

// GJK over the particles' vertex replicas (gjk.cpp:25-100 with the core of
// gjk.cpp:418-460), one wave per particle pair: the support searches are
// lane-parallel (first index of the strict extremum after a wave
// reduction); the simplex arithmetic is uniform across the wave.  At most one
// contact per pair, when |P-Q| <= epsA + epsB; triA = triB = -1.
__device__ __forceinline__ int wave_extremum(double bd, int bi, bool maximal) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double od = __shfl_xor(bd, o);
    const int oi = __shfl_xor(bi, o);
    const bool better = oi >= 0 && (bi < 0 || (maximal ? od > bd : od < bd) ||
                                    (od == bd && oi < bi));
    bd = better ? od : bd;
    bi = better ? oi : bi;
  }
  return bi;
}

__global__ __launch_bounds__(256) void k_gjk(DetectArgs g) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= g.npairs) return;
  unsigned long long* ts = g.tile_stats + static_cast<long long>(k) * kTileStats;
  if (g.npairs_dev && k >= *g.npairs_dev) {
    if (lane == 0) g.tile_count[k] = 0;
    if (lane < kTileStats) ts[lane] = 0;
    return;
  }
  const Pair p = load_pair(g, k);
  const int na = 3 * p.nA, nb = 3 * p.nB, ba = 3 * p.oA, bb = 3 * p.oB;
  auto pa = [&](int i) { return V3<double>{g.x[ba + i], g.y[ba + i], g.z[ba + i]}; };
  auto pb = [&](int i) { return V3<double>{g.x[bb + i], g.y[bb + i], g.z[bb + i]}; };
  auto min_a = [&](const double* v) {
    double bd = 1.7976931348623157e308;
    int bi = -1;
    for (int i = lane; i < na; i += 64) {
      const double c[3] = {g.x[ba + i], g.y[ba + i], g.z[ba + i]};
      const double d = gdot(c, v);
      if (d < bd) {
        bd = d;
        bi = i;
      }
    }
    return wave_extremum(bd, bi, false);
  };
  auto max_b = [&](const double* v) {
    double bd = -1.7976931348623157e308;
    int bi = -1;
    for (int i = lane; i < nb; i += 64) {
      const double c[3] = {g.x[bb + i], g.y[bb + i], g.z[bb + i]};
      const double d = gdot(c, v);
      if (d > bd) {
        bd = d;
        bi = i;
      }
    }
    return wave_extremum(bd, bi, true);
  };
  double P[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
  bool ok = na > 0 && nb > 0;
  if (ok) gjk_core(na, nb, pa, pb, min_a, max_b, P, Q, &ok);
  const V3<double> vP = {P[0], P[1], P[2]}, vQ = {Q[0], Q[1], Q[2]};
  const bool emit = ok && pq_dist(vP, vQ) <= p.margin;
  if (lane == 0) {
    if (emit) {
      const unsigned long long slot = atomicAdd(g.counter, 1ull);
      if (static_cast<long long>(slot) < g.cap) {
        dh_contact c;
        make_contact(p, k, -1, -1, vP, vQ, &c);
        c.pad_ = 0;
        g.temp[slot] = c;
      }
    }
    g.tile_count[k] = emit ? 1 : 0;
  }
  if (lane < kTileStats) {
    ts[lane] = lane == TS_CHECKS ? static_cast<unsigned long long>(p.nA) * p.nB
             : lane == TS_EVALUATED ? 1ull : 0ull;
  }
}

// exclusive scan of per-pair counts (two levels, 1024 per block)
// live (may be null): the scan covers the first min(n, *live) entries only.
__device__ __forceinline__ int scan_len(int n, const unsigned long long* live) {
  return live && *live < static_cast<unsigned long long>(n) ? static_cast<int>(*live) : n;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(const int* __restrict__ in, int n,
                                                      long long* __restrict__ out,
                                                      long long* __restrict__ block_sums,
                                                      const unsigned long long* __restrict__ live) {
  __shared__ long long buf[1024];
  n = scan_len(n, live);
  if (static_cast<long long>(blockIdx.x) * 1024 >= n) {  // past the live entries
    if (threadIdx.x == 0) block_sums[blockIdx.x] = 0;
    return;
  }
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const long long v = (i < n) ? in[i] : 0;
  buf[threadIdx.x] = v;
  __syncthreads();
  for (int s = 1; s < 1024; s <<= 1) {
    const long long t = (threadIdx.x >= s) ? buf[threadIdx.x - s] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = buf[threadIdx.x] - v;
  if (threadIdx.x == 1023) block_sums[blockIdx.x] = buf[1023];
}

// The whole scan in one workgroup when n <= 1,024 (the small scenes' pair
// and particle counts): offs[0..n) exclusive, offs[n] the total.
__global__ __launch_bounds__(1024) void k_scan_single(const int* __restrict__ in, int n,
                                                      long long* __restrict__ out,
                                                      const unsigned long long* __restrict__ live) {
  __shared__ long long buf[1024];
  n = scan_len(n, live);
  const int i = threadIdx.x;
  const long long v = i < n ? in[i] : 0;
  buf[i] = v;
  __syncthreads();
  for (int s = 1; s < 1024; s <<= 1) {
    const long long t = i >= s ? buf[i - s] : 0;
    __syncthreads();
    buf[i] += t;
    __syncthreads();
  }
  if (i < n) out[i] = buf[i] - v;
  if (i == 1023) out[n] = buf[1023];
}

// Exclusive scan of the block sums in place, by one workgroup: 1,024 at a
// time with a running carry.
__global__ __launch_bounds__(1024) void k_scan_sums(long long* __restrict__ sums, int nb) {
  __shared__ long long buf[1024];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const long long v = i < nb ? sums[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {
      const long long t = (threadIdx.x >= s) ? buf[threadIdx.x - s] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    const long long c = carry;
    if (i < nb) sums[i] = c + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c + buf[1023];
    __syncthreads();
  }
}

// Adds the block offsets; the last element's thread also writes the total
// out[n] (out[0] = 0 when n = 0).
__global__ __launch_bounds__(1024) void k_scan_add(long long* __restrict__ out, int n,
                                                   const long long* __restrict__ sums,
                                                   const int* __restrict__ in,
                                                   const unsigned long long* __restrict__ live) {
  n = scan_len(n, live);
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) {
    const long long o = out[i] + sums[blockIdx.x];
    out[i] = o;
    if (i == n - 1) out[n] = o + in[i];
  } else if (n == 0 && i == 0) {
    out[0] = 0;
  }
}

// k_scan_add for at most 1,024 blocks, without the block-sum scan: each
// block first adds up the raw sums of the blocks before it.
__global__ __launch_bounds__(1024) void k_scan_add_prefix(long long* __restrict__ out, int n,
                                                          const long long* __restrict__ sums,
                                                          const int* __restrict__ in,
                                                          const unsigned long long* __restrict__ live) {
  __shared__ long long part[16];
  n = scan_len(n, live);
  const int b = blockIdx.x, tid = threadIdx.x;
  if (static_cast<long long>(b) * 1024 >= n && !(n == 0 && b == 0)) return;
  long long v = tid < b ? sums[tid] : 0;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((tid & 63) == 0) part[tid >> 6] = v;
  __syncthreads();
  long long base = 0;
  for (int w = 0; w < 16; ++w) base += part[w];
  const int i = b * 1024 + tid;
  if (i < n) {
    const long long o = out[i] + base;
    out[i] = o;
    if (i == n - 1) out[n] = o + in[i];
  } else if (n == 0 && i == 0) {
    out[0] = 0;
  }
}

// Per-pair contact counts from the per-tile counts.
__global__ __launch_bounds__(256) void k_tile_sums(const int* __restrict__ tile_count, int tiles,
                                                   int npairs, int* __restrict__ counts) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= npairs) return;
  int c = 0;
  for (int t = 0; t < tiles; ++t) c += tile_count[static_cast<long long>(k) * tiles + t];
  counts[k] = c;
}

// Contacts into pair order: pair offset + the counts of the pair's earlier
// tiles + the rank within the tile.
__global__ __launch_bounds__(256) void k_scatter(const dh_contact* __restrict__ temp, long long n,
                                                 const long long* __restrict__ offs,
                                                 const int* __restrict__ tile_count, int tiles,
                                                 dh_contact* __restrict__ out,
                                                 const unsigned long long* __restrict__ n_dev) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  // n is the capacity (of temp and out) when the count is read from n_dev
  if (i >= n || (n_dev && static_cast<unsigned long long>(i) >= *n_dev)) return;
  dh_contact c = temp[i];
  const int pair = c.pair / tiles;
  long long dst = offs[pair] + c.pad_;
  for (int t = pair * tiles; t < c.pair; ++t) dst += tile_count[t];
  if (dst >= n) return;  // more contacts than the capacity (reported by the caller)
  c.pair = pair;
  c.pad_ = 0;
  out[dst] = c;
}

template <bool SCREEN>
__device__ __forceinline__ void queue_solve_entry(const DetectArgs& g, long long e,
                                                  int* __restrict__ flags,
                                                  double* __restrict__ pq) {
  const int2 en = g.queue[e];
  if (en.x < 0) {
    flags[e] = 0;
    return;
  }
  const int k = en.x / g.tiles;
  const int pa = g.pairs[2 * k], pb = g.pairs[2 * k + 1];
  const int iA = static_cast<int>(static_cast<unsigned>(en.y) & 0xffffu);
  const int iB = static_cast<int>(static_cast<unsigned>(en.y) >> 16);
  const Tri<double> A = load_tri(g, g.tri_offset[pa] + iA);
  const Tri<double> B = load_tri(g, g.tri_offset[pb] + iB);
  const double margin = g.eps[pa] + g.eps[pb];
  if (axes_apart(A, B, margin) || (SCREEN && screened_far(A, B, margin))) {
    flags[e] = 0;
    return;
  }
  V3<double> P, Q;
  bf_solver(A, B, P, Q);
  const bool emit = pq_dist(P, Q) <= margin;
  flags[e] = emit ? 1 : 0;
  if (emit) {
    double* o = pq + 6 * e;
    o[0] = P.x; o[1] = P.y; o[2] = P.z;
    o[3] = Q.x; o[4] = Q.y; o[5] = Q.z;
  }
}


// Survivor queue, pass 2: a lane per queued triangle pair at a time
// (bfSolver, the contact test of bf.cpp:125); flags[e] = 1 and P, Q kept when
// it emits.  The entries are the first n, or, with need_dev (a captured step
// launched for the capacity n), the first min(*need_dev, n); the later passes
// stop at the same count.
template <bool SCREEN>
__global__ __launch_bounds__(256) void k_queue_solve(DetectArgs g, long long n,
                                                     const unsigned long long* __restrict__ need_dev,
                                                     int* __restrict__ flags,
                                                     double* __restrict__ pq) {
  // the live entries (flags past them are not written: the scan stops there)
  const long long live =
      need_dev && static_cast<long long>(*need_dev) < n ? static_cast<long long>(*need_dev) : n;
  for (long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; e < live;
       e += static_cast<long long>(gridDim.x) * 256)
    queue_solve_entry<SCREEN>(g, e, flags, pq);
}

// Pass 3: the queue's contacts take one range of the contact slots.
__device__ __forceinline__ long long queue_live(long long n,
                                                const unsigned long long* need_dev) {
  return need_dev && static_cast<long long>(*need_dev) < n ? static_cast<long long>(*need_dev) : n;
}

__global__ void k_queue_reserve(DetectArgs g, const long long* __restrict__ pos, long long n,
                                const unsigned long long* __restrict__ need_dev,
                                unsigned long long* __restrict__ base) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *base = atomicAdd(g.counter, static_cast<unsigned long long>(pos[queue_live(n, need_dev)]));
}

// Pass 4: contacts of the emitting entries, tagged (item, rank within the
// item): the item's entries are contiguous and in row-major order, so the
// rank is the number of emitting entries before it in the item's range.
__device__ __forceinline__ void queue_emit_entry(const DetectArgs& g, long long e,
                                                 const long long* __restrict__ pos,
                                                 const double* __restrict__ pq,
                                                 const unsigned long long* __restrict__ base) {
  const int2 en = g.queue[e];
  const int item = en.x;
  const int k = item / g.tiles;
  const Pair p = load_pair(g, k);
  const int iA = static_cast<int>(static_cast<unsigned>(en.y) & 0xffffu);
  const int iB = static_cast<int>(static_cast<unsigned>(en.y) >> 16);
  const double* o = pq + 6 * e;
  const V3<double> P = {o[0], o[1], o[2]}, Q = {o[3], o[4], o[5]};
  const unsigned long long slot = *base + static_cast<unsigned long long>(pos[e]);
  if (slot < static_cast<unsigned long long>(g.cap)) {
    dh_contact c;
    make_contact(p, k, iA, iB, P, Q, &c);
    c.pair = item;
    c.pad_ = static_cast<int>(pos[e] - pos[g.qbase[item]]);
    g.temp[slot] = c;
  }
  atomicAdd(&g.tile_count[item], 1);
}

__global__ __launch_bounds__(256) void k_queue_emit(DetectArgs g, long long n,
                                                    const unsigned long long* __restrict__ need_dev,
                                                    const int* __restrict__ flags,
                                                    const long long* __restrict__ pos,
                                                    const double* __restrict__ pq,
                                                    const unsigned long long* __restrict__ base) {
  const long long live = queue_live(n, need_dev);
  for (long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; e < live;
       e += static_cast<long long>(gridDim.x) * 256)
    if (flags[e]) queue_emit_entry(g, e, pos, pq, base);
}

int queue_learned(const int2* queue, const int* pairs, int tiles, const int* tri_offset,
                  const double* x, const double* y, const double* z, const double* eps,
                  long long n, const unsigned long long* need_dev, int* flags, double* pq,
                  hipStream_t st);
int queue_learned_reserve(long long n);

struct DetectScratch {
  DevBuf temp, counter, stats, error, counts, offs, sums, tiles, tstats;
  DevBuf queue, qcounter, qbase, qflags, qpos, qpq, qres, deferred, ndeferred, iskip, tmax;
  long long qcap = 0;  // queue entries of the next eager DH_BF call
};
static DetectScratch g_scratch;

// The largest triangle count of a particle, per geometry (tri_offset, n,
// ntri): k_chunk_cull runs only on scenes with particles of more than
// kRowBlock triangles.  A stale entry changes the speed, never the result
// (the pre-pass is exact, and without it detect_item does the same tests).
struct TriMax {
  const int* tri_offset = nullptr;
  int n = -1, ntri = -1, max_tris = 0;
};
static TriMax g_trimax;

__global__ __launch_bounds__(256) void k_max_tris(const int* __restrict__ off, int n,
                                                  int* __restrict__ out) {
  int m = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    m = max(m, off[i + 1] - off[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

// With sync, computes and caches a geometry not cached (one read-back);
// without (inside a capture), answers 0 for it.
static int particle_max_tris(const dh_geometry* g, hipStream_t st, bool sync) {
  if (g_trimax.tri_offset == g->tri_offset && g_trimax.n == g->nparticles &&
      g_trimax.ntri == g->ntri_total)
    return g_trimax.max_tris;
  if (!sync || g->nparticles <= 0) return 0;
  if (g_scratch.tmax.grow(sizeof(int))) return 0;
  int m = 0;
  if (hipMemsetAsync(g_scratch.tmax.ptr, 0, sizeof(int), st) != hipSuccess) return 0;
  const unsigned grid = grid_for(g->nparticles, 256) < 1024 ? grid_for(g->nparticles, 256) : 1024;
  hipLaunchKernelGGL(k_max_tris, grid, 256, 0, st, g->tri_offset, g->nparticles,
                     g_scratch.tmax.as<int>());
  if (hipMemcpyAsync(&m, g_scratch.tmax.ptr, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return 0;
  g_trimax = {g->tri_offset, g->nparticles, g->ntri_total, m};
  return m;
}

// Before a capture of a step on this geometry: fills the cache above.
int detect_prepare(const dh_geometry* geom, hipStream_t st) {
  particle_max_tris(geom, st, true);
  return DH_OK;
}

// Scratch of scan_counts for up to n entries (before a capture).
int scan_reserve(int n) {
  const int nb = (n + 1023) / 1024 > 0 ? (n + 1023) / 1024 : 1;
  return g_scratch.sums.grow(sizeof(long long) * (nb + 1));
}

// Exclusive scan of counts[0..m) into offs[0..m), offs[m] the total, where
// m = n, or min(n, *live) when live is given (read when the kernels run);
// `sums` holds scan_sums_len(n) block sums (unused when n <= 1,024).
static int scan_counts_upto(const int* counts, int n, const unsigned long long* live,
                            long long* offs, long long* sums, hipStream_t st) {
  if (n <= 1024) {
    hipLaunchKernelGGL(k_scan_single, 1, 1024, 0, st, counts, n, offs, live);
    DH_HIP_CHECK(hipGetLastError());
    return DH_OK;
  }
  const int nb = (n + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_blocks, nb, 1024, 0, st, counts, n, offs, sums, live);
  if (nb <= 1024) {  // two launches: the add pass sums the block sums before it
    hipLaunchKernelGGL(k_scan_add_prefix, nb, 1024, 0, st, offs, n, sums, counts, live);
  } else {
    hipLaunchKernelGGL(k_scan_sums, 1, 1024, 0, st, sums, nb);
    hipLaunchKernelGGL(k_scan_add, nb, 1024, 0, st, offs, n, sums, counts, live);
  }
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

// offs[0..n]: exclusive prefix sums of counts, offs[n] = total.
int scan_counts_with(const int* counts, int n, long long* offs, long long* sums,
                     hipStream_t st) {
  return scan_counts_upto(counts, n, nullptr, offs, sums, st);
}

int scan_sums_len(int n) { return (n + 1023) / 1024 > 0 ? (n + 1023) / 1024 + 1 : 2; }

// The same with the library's own block-sum scratch.
int scan_counts(const int* counts, int n, long long* offs, hipStream_t st) {
  if (n > 1024) {
    int rc = g_scratch.sums.grow(sizeof(long long) * scan_sums_len(n));
    if (rc) return rc;
  }
  return scan_counts_with(counts, n, offs, g_scratch.sums.as<long long>(), st);
}

// Workgroups of a detection launch over `items`: one per item, or the
// persistent grid of detect_wg_per_cu workgroups per CU.
static unsigned detect_grid(long long items) {
  const int per_cu = tuning().detect_wg_per_cu;
  long long g = items;
  if (per_cu > 0 && g > static_cast<long long>(device_cus()) * per_cu)
    g = static_cast<long long>(device_cus()) * per_cu;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

template <int M>
static void launch_detect(unsigned items, size_t shm, hipStream_t st, const DetectArgs& a) {
  if (M == DH_BF && a.queue && a.wave_items) {
    hipLaunchKernelGGL(k_lane_cull, items > 0 ? grid_for(items, 256) : 1u, 256, 0, st, a,
                       static_cast<int>(items));
    // at most the workgroups one CU holds at once (3 waves per SIMD at the
    // kernel's 161 VGPRs, 2 waves per workgroup): no partial second round
    const long long wg = (static_cast<long long>(items) + 1) / 2;
    const long long cap = static_cast<long long>(device_cus()) * 6;
    hipLaunchKernelGGL(k_detect_wave, static_cast<unsigned>(wg < cap ? (wg > 0 ? wg : 1) : cap),
                       128, 0, st, a, static_cast<int>(items));
  }
  if (M == DH_BF && a.queue && a.wave_items) {
    const long long wg = (static_cast<long long>(items) + 63) / 64;
    const long long cap = static_cast<long long>(device_cus()) * 32;
    hipLaunchKernelGGL(k_detect_big, static_cast<unsigned>(wg < cap ? (wg > 0 ? wg : 1) : cap),
                       256, shm, st, a, static_cast<int>(items));
  } else if (M == DH_BF && a.queue) {
    hipLaunchKernelGGL((k_detect<DH_BF, true>), detect_grid(items), 256, shm, st, a,
                       static_cast<int>(items));
  } else {
    if (M == DH_BF && a.item_skip) {
      const long long wg = (static_cast<long long>(items) + 3) / 4;
      const long long cap = static_cast<long long>(device_cus()) * 32;
      hipLaunchKernelGGL(k_chunk_cull, static_cast<unsigned>(wg < cap ? (wg > 0 ? wg : 1) : cap),
                         256, 0, st, a, static_cast<int>(items),
                         const_cast<unsigned char*>(a.item_skip));
    }
    hipLaunchKernelGGL((k_detect<M, false>), detect_grid(items), 256, shm, st, a,
                       static_cast<int>(items));
  }
}

static void launch_tile_stats(const unsigned long long* ts, long long ntiles,
                              unsigned long long* stats, hipStream_t st) {
  if (ntiles <= 0) return;
  long long grid = (ntiles + 255) / 256;
  if (grid > 256) grid = 256;
  hipLaunchKernelGGL(k_tile_stats, static_cast<unsigned>(grid), 256, 0, st, ts, ntiles, stats);
}

// Row ranges per particle pair: enough workgroups to fill the 256 CUs
// several times over when the pairs are few (detect_tile_target workgroups
// in all, at most detect_max_tiles per pair), one per pair otherwise.
static int tiles_for(long long npairs, int method) {
  if (method == DH_GJK || npairs <= 0) return 1;
  const long long target = tuning().detect_tile_target, cap = tuning().detect_max_tiles;
  long long t = (target + npairs - 1) / npairs;
  return static_cast<int>(t < 1 ? 1 : (t > cap ? cap : t));
}

// Buffers of the queue passes for n entries.
static int queue_reserve(long long n) {
  int rc;
  if ((rc = g_scratch.qflags.grow(sizeof(int) * (n + 1)))) return rc;
  if ((rc = g_scratch.qpos.grow(sizeof(long long) * (n + 1)))) return rc;
  if ((rc = g_scratch.qpq.grow(sizeof(double) * 6 * n))) return rc;
  if ((rc = g_scratch.qres.grow(sizeof(unsigned long long)))) return rc;
  return scan_reserve(static_cast<int>(n));
}

// Solve, scan, reserve and emit over n queue entries (live: all of them, or
// min(*need_dev, n) in a captured step).
static int queue_passes(const DetectArgs& a, long long n, const unsigned long long* need_dev,
                        bool learned, hipStream_t st) {
  // a captured step's passes are launched for the capacity: a persistent grid
  // strides over the live entries
  long long g = grid_for(n, 256);
  const long long gcap = static_cast<long long>(device_cus()) * 16;
  if (need_dev && g > gcap) g = gcap;
  const unsigned grid = static_cast<unsigned>(g > 0 ? g : 1);
  int rc;
  if (learned) {
    if ((rc = queue_learned(a.queue, a.pairs, a.tiles, a.tri_offset, a.x, a.y, a.z, a.eps, n,
                            need_dev, g_scratch.qflags.as<int>(), g_scratch.qpq.as<double>(),
                            st)))
      return rc;
  } else if (a.fp32_screen)
    hipLaunchKernelGGL(k_queue_solve<true>, grid, 256, 0, st, a, n, need_dev,
                       g_scratch.qflags.as<int>(), g_scratch.qpq.as<double>());
  else
    hipLaunchKernelGGL(k_queue_solve<false>, grid, 256, 0, st, a, n, need_dev,
                       g_scratch.qflags.as<int>(), g_scratch.qpq.as<double>());
  const int ni = static_cast<int>(n);
  if (ni > 1024 && (rc = g_scratch.sums.grow(sizeof(long long) * scan_sums_len(ni)))) return rc;
  if ((rc = scan_counts_upto(g_scratch.qflags.as<int>(), static_cast<int>(n), need_dev,
                             g_scratch.qpos.as<long long>(), g_scratch.sums.as<long long>(), st)))
    return rc;
  hipLaunchKernelGGL(k_queue_reserve, 1, 64, 0, st, a, g_scratch.qpos.as<long long>(), n, need_dev,
                     g_scratch.qres.as<unsigned long long>());
  hipLaunchKernelGGL(k_queue_emit, grid, 256, 0, st, a, n, need_dev, g_scratch.qflags.as<int>(),
                     g_scratch.qpos.as<long long>(), g_scratch.qpq.as<double>(),
                     g_scratch.qres.as<unsigned long long>());
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

// Passes 2-4 of the survivor queue after the detection launch: the queued
// triangle pairs solved one per lane, an exclusive scan of the emitting
// flags, one reservation of contact slots, the contacts.  Reads the queue
// length back (one synchronisation); the next call's queue holds this
// call's need.
static int solve_queue(const DetectArgs& a, bool learned, hipStream_t st) {
  unsigned long long need = 0;
  unsigned ndef = 0;
  DH_HIP_CHECK(hipMemcpyAsync(&need, a.qcounter, sizeof(need), hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipMemcpyAsync(&ndef, a.ndeferred, sizeof(ndef), hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  if (ndef > 0) {
    // the items the mid-phase launch could not queue, in full
    DetectArgs f = a;
    f.queue = nullptr;
    f.item_list = a.deferred;
    f.nitems_dev = nullptr;
    hipLaunchKernelGGL((k_detect<DH_BF, false>), detect_grid(ndef), 256, 0, st, f,
                       static_cast<int>(ndef));
  }
  if (static_cast<long long>(need) > a.qcap)
    g_scratch.qcap = static_cast<long long>(need) + static_cast<long long>(need) / 4;
  const long long live = static_cast<long long>(need) < a.qcap ? static_cast<long long>(need) : a.qcap;
  if (live == 0) return DH_OK;
  int rc;
  if ((rc = queue_reserve(live))) return rc;
  return queue_passes(a, live, nullptr, learned, st);
}

int detect_pairs_impl(int method, const dh_geometry* geom, const int* pairs, int npairs,
                      dh_contact* out, long long cap, long long* n_total, int* pair_count_user,
                      long long* offs_out, dh_stats* stats, hipStream_t st,
                      const double* pbox, const double* bbox, const double* psph) {
  int rc;
  // DH_BF_LEARNED: DH_BF with the survivor queue always on, its pairs
  // through the learned decision
  const bool learned = method == DH_BF_LEARNED;
  if (learned) method = DH_BF;
  const long long temp_cap = cap > 0 ? cap : 1;
  if ((rc = g_scratch.temp.grow(sizeof(dh_contact) * temp_cap))) return rc;
  if ((rc = g_scratch.counter.grow(sizeof(unsigned long long)))) return rc;
  if ((rc = g_scratch.stats.grow(sizeof(unsigned long long) * ST_COUNT))) return rc;
  if ((rc = g_scratch.error.grow(sizeof(int)))) return rc;
  if ((rc = g_scratch.counts.grow(sizeof(int) * (npairs + 1)))) return rc;
  if ((rc = g_scratch.offs.grow(sizeof(long long) * (npairs + 1)))) return rc;
  DH_HIP_CHECK(hipMemsetAsync(g_scratch.counter.ptr, 0, sizeof(unsigned long long), st));
  DH_HIP_CHECK(hipMemsetAsync(g_scratch.stats.ptr, 0, sizeof(unsigned long long) * ST_COUNT, st));
  DH_HIP_CHECK(hipMemsetAsync(g_scratch.error.ptr, 0, sizeof(int), st));
  int* counts = pair_count_user ? pair_count_user : g_scratch.counts.as<int>();
  const int tiles = tiles_for(npairs, method);
  int* tile_count = counts;
  if (tiles > 1) {
    if ((rc = g_scratch.tiles.grow(sizeof(int) * static_cast<size_t>(npairs) * tiles))) return rc;
    tile_count = g_scratch.tiles.as<int>();
  }
  DetectArgs a;
  a.x = geom->x;
  a.y = geom->y;
  a.z = geom->z;
  a.tri_offset = geom->tri_offset;
  a.eps = geom->eps;
  a.friction = geom->friction;
  a.id = geom->id;
  a.pairs = pairs;
  a.npairs = npairs;
  a.temp = g_scratch.temp.as<dh_contact>();
  a.cap = cap;
  a.counter = g_scratch.counter.as<unsigned long long>();
  a.tiles = tiles;
  a.tile_count = tile_count;
  const long long ntiles = static_cast<long long>(npairs) * tiles;
  if ((rc = g_scratch.tstats.grow(sizeof(unsigned long long) * kTileStats * ntiles))) return rc;
  a.tile_stats = g_scratch.tstats.as<unsigned long long>();
  a.stats = g_scratch.stats.as<unsigned long long>();
  a.error = g_scratch.error.as<int>();
  a.npairs_dev = nullptr;
  a.pbox = pbox;
  a.psph = pbox ? psph : nullptr;
  a.fp32_screen = tuning().detect_fp32_screen;
  a.wave_items = 0;
  a.item_skip = nullptr;
  a.bbox = bbox;
  a.queue = nullptr;
  a.qcap = 0;
  a.qcounter = nullptr;
  a.qbase = nullptr;
  a.deferred = nullptr;
  a.ndeferred = nullptr;
  a.biglist = nullptr;
  a.nbig = nullptr;
  a.wlist = nullptr;
  a.nwlist = nullptr;
  a.item_list = nullptr;
  a.nitems_dev = nullptr;
  const long long nitems = static_cast<long long>(npairs) * tiles;
  // the queue costs a read-back and four launches: for many items only
  if (method == DH_BF && (learned || nitems >= tuning().detect_queue_min_items)) {
    // survivor queue: sized from the last call's need, at least 16 entries
    // per item (an overflowing item solves its own survivors)
    long long qcap = g_scratch.qcap > 16 * nitems + 65536 ? g_scratch.qcap : 16 * nitems + 65536;
    if (qcap > (1ll << 30)) qcap = 1ll << 30;
    if ((rc = g_scratch.queue.grow(sizeof(int2) * qcap))) return rc;
    if ((rc = g_scratch.qcounter.grow(sizeof(unsigned long long)))) return rc;
    if ((rc = g_scratch.qbase.grow(sizeof(long long) * (nitems + 1)))) return rc;
    if ((rc = g_scratch.deferred.grow(sizeof(int) * 3 * (nitems + 1)))) return rc;
    if ((rc = g_scratch.ndeferred.grow(sizeof(unsigned) * 3))) return rc;
    DH_HIP_CHECK(hipMemsetAsync(g_scratch.qcounter.ptr, 0, sizeof(unsigned long long), st));
    DH_HIP_CHECK(hipMemsetAsync(g_scratch.ndeferred.ptr, 0, sizeof(unsigned) * 3, st));
    a.deferred = g_scratch.deferred.as<int>();
    a.ndeferred = g_scratch.ndeferred.as<unsigned>();
    a.biglist = a.deferred + (nitems + 1);
    a.nbig = a.ndeferred + 1;
    a.wlist = a.deferred + 2 * (nitems + 1);
    a.nwlist = a.ndeferred + 2;
    a.wave_items = 1;
    a.queue = g_scratch.queue.as<int2>();
    a.qcap = qcap;
    a.qcounter = g_scratch.qcounter.as<unsigned long long>();
    a.qbase = g_scratch.qbase.as<long long>();
  }
  const size_t shm_rows = 2 * kMaxRowTriangles * sizeof(int) + 8 + kMaxRowTriangles * 8;
  const unsigned items = static_cast<unsigned>(static_cast<long long>(npairs) * tiles);
  if (method == DH_BF && !a.queue && pbox && bbox &&
      particle_max_tris(geom, st, true) > kRowBlock) {
    if ((rc = g_scratch.iskip.grow(items > 0 ? items : 1))) return rc;
    a.item_skip = g_scratch.iskip.as<unsigned char>();
  }
  switch (method) {
    case DH_BF: launch_detect<DH_BF>(items, 0, st, a); break;
    case DH_PENALTY: launch_detect<DH_PENALTY>(items, 0, st, a); break;
    case DH_PENALTY_STAT: launch_detect<DH_PENALTY_STAT>(items, 0, st, a); break;
    case DH_HYBRID_BATCH: launch_detect<DH_HYBRID_BATCH>(items, shm_rows, st, a); break;
    case DH_HYBRID_TRIANGLE: launch_detect<DH_HYBRID_TRIANGLE>(items, shm_rows, st, a); break;
    case DH_HYBRID_BATCH_STAT: launch_detect<DH_HYBRID_BATCH_STAT>(items, shm_rows, st, a); break;
    case DH_HYBRID_TRIANGLE_STAT:
      launch_detect<DH_HYBRID_TRIANGLE_STAT>(items, shm_rows, st, a);
      break;
    case DH_GJK:
      hipLaunchKernelGGL(k_gjk, (items + 3) / 4, 256, 0, st, a);
      break;
    default:
      return set_error(DH_ERR_UNSUPPORTED, "dh_detect_pairs: method %d", method);
  }
  DH_HIP_CHECK(hipGetLastError());
  if (a.queue && (rc = solve_queue(a, learned, st))) return rc;
  launch_tile_stats(a.tile_stats, ntiles, a.stats, st);
  unsigned long long total = 0;
  int err = 0;
  unsigned long long st_host[ST_COUNT];
  DH_HIP_CHECK(hipMemcpyAsync(&total, g_scratch.counter.ptr, sizeof(total),
                              hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipMemcpyAsync(&err, g_scratch.error.ptr, sizeof(err), hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipMemcpyAsync(st_host, g_scratch.stats.ptr, sizeof(st_host),
                              hipMemcpyDeviceToHost, st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  if (err == DE_REF_UB)
    return set_error(DH_ERR_REF_UB,
                     "hybridWithPerTriangleFallBack reads row iA < nTriB of A: nTriB > nTriA is "
                     "out of bounds in the reference (hybrid.cpp:130)");
  if (err == DE_TOO_LONG)
    return set_error(DH_ERR_ARG, "more than %d triangles in a particle (reference row limit)",
                     kMaxRowTriangles);
  *n_total = static_cast<long long>(total);
  if (stats) {
    for (int i = 0; i < 17; ++i) stats->newton_histogram[i] += st_host[ST_HIST + i];
    stats->penalty_fails += st_host[ST_PENALTY_FAILS];
    stats->batch_fails += st_host[ST_BATCH_FAILS];
    stats->batch_size += st_host[ST_BATCH_SIZE];
    stats->triangle_checks += st_host[ST_CHECKS];
    stats->pairs_evaluated += st_host[ST_EVALUATED];
    stats->fallback_rows += st_host[ST_FALLBACK_ROWS];
    stats->pairs_culled += st_host[ST_CULLED];
  }
  if (static_cast<long long>(total) > cap) return DH_ERR_CAPACITY;
  long long* offs = offs_out ? offs_out : g_scratch.offs.as<long long>();
  if (tiles > 1)
    hipLaunchKernelGGL(k_tile_sums, grid_for(npairs, 256), 256, 0, st, tile_count, tiles, npairs,
                       counts);
  if ((rc = scan_counts(counts, npairs, offs, st))) return rc;
  if (total > 0) {
    hipLaunchKernelGGL(k_scatter, grid_for(static_cast<int64_t>(total), 256), 256, 0, st,
                       g_scratch.temp.as<dh_contact>(), static_cast<long long>(total), offs,
                       tile_count, tiles, out, nullptr);
    DH_HIP_CHECK(hipGetLastError());
  }
  return DH_OK;
}

int detect_reserve(int pair_cap, long long contact_cap, int method);

// Tiles of the captured step, sized for half the pair capacity live.
static int graph_tiles(int pair_cap) { return tiles_for(pair_cap / 2 > 0 ? pair_cap / 2 : 1, DH_BF); }

// Survivor-queue entries of a captured DH_BF step (0: no queue).
static long long graph_qcap(int method, int pair_cap) {
  const long long items = static_cast<long long>(pair_cap) * graph_tiles(pair_cap);
  if (method == DH_BF_LEARNED) method = DH_BF;
  else if (method != DH_BF || items < tuning().detect_queue_min_items) return 0;
  const long long q = 16 * items + 65536;
  return q < (1ll << 30) ? q : (1ll << 30);
}

// Detection for a captured step: launches for `pair_cap` pairs of which
// *np_dev are live, contacts into `out` (capacity contact_cap), per-pair
// counts and offsets; no host synchronisation.  Counters stay on the device:
// dev_counters()[0] = contacts found (may exceed contact_cap), error flag.
int detect_pairs_graph(int method, const dh_geometry* geom, const int* pairs, int pair_cap,
                       const long long* np_dev, dh_contact* out, long long contact_cap,
                       int* counts, long long* offs, hipStream_t st, const double* pbox,
                       const double* bbox, const double* psph) {
  int rc;
  if ((rc = detect_reserve(pair_cap, contact_cap, method))) return rc;
  const bool learned = method == DH_BF_LEARNED;
  const long long qcap = graph_qcap(method, pair_cap);
  if (learned) method = DH_BF;
  zero_async(ZeroRegions{{g_scratch.counter.as<unsigned>(), g_scratch.stats.as<unsigned>(),
                          g_scratch.error.as<unsigned>(),
                          qcap > 0 ? g_scratch.qcounter.as<unsigned>() : nullptr},
                         {2, 2 * ST_COUNT, 1, 2}},
             st);
  DetectArgs a;
  a.x = geom->x;
  a.y = geom->y;
  a.z = geom->z;
  a.tri_offset = geom->tri_offset;
  a.eps = geom->eps;
  a.friction = geom->friction;
  a.id = geom->id;
  a.pairs = pairs;
  a.npairs = pair_cap;
  a.temp = g_scratch.temp.as<dh_contact>();
  a.cap = contact_cap;
  a.counter = g_scratch.counter.as<unsigned long long>();
  const int tiles = graph_tiles(pair_cap);
  a.tiles = tiles;
  a.tile_count = tiles > 1 ? g_scratch.tiles.as<int>() : counts;
  const long long ntiles = static_cast<long long>(pair_cap) * tiles;
  a.tile_stats = g_scratch.tstats.as<unsigned long long>();
  a.stats = g_scratch.stats.as<unsigned long long>();
  a.error = g_scratch.error.as<int>();
  a.npairs_dev = np_dev;
  a.pbox = pbox;
  a.psph = pbox ? psph : nullptr;
  a.fp32_screen = tuning().detect_fp32_screen;
  a.wave_items = 0;
  a.item_skip = nullptr;
  a.bbox = bbox;
  a.queue = nullptr;  // the queue needs the survivor count on the host
  a.qcap = 0;
  a.qcounter = nullptr;
  a.qbase = nullptr;
  a.deferred = nullptr;
  a.ndeferred = nullptr;
  a.biglist = nullptr;
  a.nbig = nullptr;
  a.wlist = nullptr;
  a.nwlist = nullptr;
  a.item_list = nullptr;
  a.nitems_dev = nullptr;
  const unsigned items = static_cast<unsigned>(static_cast<long long>(pair_cap) * tiles);
  // the chunk pre-pass on the geometries detect_prepare found walls in
  if (method == DH_BF && qcap == 0 && pbox && bbox &&
      particle_max_tris(geom, st, false) > kRowBlock)
    a.item_skip = g_scratch.iskip.as<unsigned char>();
  if (qcap > 0) {
    // the survivor queue with its counts kept on the device: every pass is
    // launched for the capacity and reads the live length
    fill_async(g_scratch.ndeferred.as<unsigned>(), 0u, 3, st);
    a.queue = g_scratch.queue.as<int2>();
    a.qcap = qcap;
    a.qcounter = g_scratch.qcounter.as<unsigned long long>();
    a.qbase = g_scratch.qbase.as<long long>();
    a.deferred = g_scratch.deferred.as<int>();
    a.ndeferred = g_scratch.ndeferred.as<unsigned>();
    a.biglist = a.deferred + (items + 1);
    a.nbig = a.ndeferred + 1;
    a.wlist = a.deferred + 2 * (items + 1);
    a.nwlist = a.ndeferred + 2;
    a.wave_items = 1;
  }
  switch (method) {
    case DH_BF: launch_detect<DH_BF>(items, 0, st, a); break;
    case DH_PENALTY: launch_detect<DH_PENALTY>(items, 0, st, a); break;
    default:
      return set_error(DH_ERR_UNSUPPORTED, "captured step: method %d (use DH_BF or DH_PENALTY)",
                       method);
  }
  if (a.queue) {
    // the items the mid-phase launch could not queue, in full (as many as
    // *ndeferred of the capacity), then the queue's passes
    DetectArgs f = a;
    f.queue = nullptr;
    f.item_list = a.deferred;
    f.nitems_dev = a.ndeferred;
    // persistent over the deferred items (usually none): a grid for the
    // capacity would cost a dispatch per item
    const unsigned dcap = static_cast<unsigned>(device_cus()) * 4;
    const unsigned dgrid = detect_grid(items) < dcap ? detect_grid(items) : dcap;
    hipLaunchKernelGGL((k_detect<DH_BF, false>), dgrid, 256, 0, st, f, static_cast<int>(items));
    if ((rc = queue_passes(a, a.qcap, a.qcounter, learned, st))) return rc;
  }
  if (pair_cap <= kTailPairs) {
    hipLaunchKernelGGL(k_detect_tail_small, 1, 1024, 0, st, a.tile_stats, ntiles, a.stats,
                       a.tile_count, tiles, pair_cap, counts, offs);
  } else {
    launch_tile_stats(a.tile_stats, ntiles, a.stats, st);
    if (tiles > 1)
      hipLaunchKernelGGL(k_tile_sums, grid_for(pair_cap, 256), 256, 0, st, a.tile_count, tiles,
                         pair_cap, counts);
    if ((rc = scan_counts(counts, pair_cap, offs, st))) return rc;
  }
  hipLaunchKernelGGL(k_scatter, grid_for(contact_cap, 256), 256, 0, st,
                     g_scratch.temp.as<dh_contact>(), contact_cap, offs, a.tile_count, tiles, out,
                     g_scratch.counter.as<unsigned long long>());
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

// Buffers of detect_pairs_graph (call before a capture: no allocation may
// happen inside one).
int detect_reserve(int pair_cap, long long contact_cap, int method) {
  int rc;
  const long long qcap = graph_qcap(method, pair_cap);
  if (qcap > 0) {
    const long long items = static_cast<long long>(pair_cap) * graph_tiles(pair_cap);
    if ((rc = g_scratch.queue.grow(sizeof(int2) * qcap))) return rc;
    if ((rc = g_scratch.qcounter.grow(sizeof(unsigned long long)))) return rc;
    if ((rc = g_scratch.qbase.grow(sizeof(long long) * (items + 1)))) return rc;
    if ((rc = g_scratch.deferred.grow(sizeof(int) * 3 * (items + 1)))) return rc;
    if ((rc = g_scratch.ndeferred.grow(sizeof(unsigned) * 3))) return rc;
    if ((rc = queue_reserve(qcap))) return rc;
    if (method == DH_BF_LEARNED && (rc = queue_learned_reserve(qcap))) return rc;
  }
  if ((rc = g_scratch.temp.grow(sizeof(dh_contact) * (contact_cap > 0 ? contact_cap : 1)))) return rc;
  if ((rc = g_scratch.counter.grow(sizeof(unsigned long long)))) return rc;
  if ((rc = g_scratch.stats.grow(sizeof(unsigned long long) * ST_COUNT))) return rc;
  if ((rc = g_scratch.error.grow(sizeof(int)))) return rc;
  const int nb = (pair_cap + 1023) / 1024 > 0 ? (pair_cap + 1023) / 1024 : 1;
  if ((rc = g_scratch.sums.grow(sizeof(long long) * (nb + 1)))) return rc;
  const int tiles = graph_tiles(pair_cap);
  if (tiles > 1 &&
      (rc = g_scratch.tiles.grow(sizeof(int) * static_cast<size_t>(pair_cap) * tiles)))
    return rc;
  if (method == DH_BF &&
      (rc = g_scratch.iskip.grow(static_cast<size_t>(pair_cap) * tiles + 1)))
    return rc;
  if ((rc = g_scratch.tstats.grow(sizeof(unsigned long long) * kTileStats *
                                  static_cast<size_t>(pair_cap) * tiles)))
    return rc;
  return DH_OK;
}

// Device addresses of the detection counters: [0] contacts found, and the
// error flag.
const unsigned long long* detect_counter_dev() { return g_scratch.counter.as<unsigned long long>(); }
const int* detect_error_dev() { return g_scratch.error.as<int>(); }
const unsigned long long* detect_stats_dev() { return g_scratch.stats.as<unsigned long long>(); }

// _batchError of the *Stat drivers: maxErr added once per row, in row order
// and pair order (hybrid.cpp:240,530).
static double batch_error_sum(double acc, double max_err, int rows) {
  for (int r = 0; r < rows; ++r) acc += max_err;
  return acc;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_exclusive_scan(const int* counts_dev, int n, int64_t* offs_dev, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (n < 0 || !offs_dev || (n > 0 && !counts_dev))
    return set_error(DH_ERR_ARG, "dh_exclusive_scan: bad arguments");
  return scan_counts(counts_dev, n, reinterpret_cast<long long*>(offs_dev), as_stream(stream));
}

int dh_detect_pairs(int method, const dh_geometry* geom, const int* pairs_dev, int npairs,
                    dh_contact* out_dev, int cap, int* n_out, int* pair_count_dev,
                    dh_stats* stats, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!geom || !n_out || npairs < 0 || cap < 0 || (npairs > 0 && !pairs_dev))
    return set_error(DH_ERR_ARG, "dh_detect_pairs: bad arguments");
  *n_out = 0;
  if (npairs == 0) return DH_OK;
  long long total = 0;
  rc = detect_pairs_impl(method, geom, pairs_dev, npairs, out_dev, cap, &total, pair_count_dev,
                         nullptr, stats, as_stream(stream), nullptr, nullptr, nullptr);
  *n_out = static_cast<int>(total);
  if (stats && (method == DH_HYBRID_BATCH_STAT || method == DH_HYBRID_TRIANGLE_STAT) &&
      (rc == DH_OK || rc == DH_ERR_CAPACITY)) {
    // sequential _batchError accumulation needs eps and row counts per pair
    std::vector<int> pr(2 * static_cast<size_t>(npairs));
    std::vector<int> off(geom->nparticles + 1);
    std::vector<double> eps(geom->nparticles);
    DH_HIP_CHECK(hipMemcpy(pr.data(), pairs_dev, sizeof(int) * pr.size(), hipMemcpyDeviceToHost));
    DH_HIP_CHECK(hipMemcpy(off.data(), geom->tri_offset, sizeof(int) * off.size(),
                           hipMemcpyDeviceToHost));
    DH_HIP_CHECK(hipMemcpy(eps.data(), geom->eps, sizeof(double) * eps.size(),
                           hipMemcpyDeviceToHost));
    double acc = stats->batch_error;
    for (int k = 0; k < npairs; ++k) {
      const int a = pr[2 * k], b = pr[2 * k + 1];
      acc = batch_error_sum(acc, (eps[a] + eps[b]) / 16, off[a + 1] - off[a]);
      (void)b;
    }
    stats->batch_error = acc;
  }
  return rc;
}

int dh_detect_pair(int method, const double* xA, const double* yA, const double* zA, int nTriA,
                   double epsA, int frictionA, int idA, const double* xB, const double* yB,
                   const double* zB, int nTriB, double epsB, int frictionB, int idB,
                   dh_contact* out, int cap, int* n_out, dh_stats* stats) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!n_out || nTriA < 0 || nTriB < 0 || cap < 0 || (cap > 0 && !out))
    return set_error(DH_ERR_ARG, "dh_detect_pair: bad arguments");
  *n_out = 0;
  if (method == DH_HYBRID_TRIANGLE && nTriB > nTriA)
    return set_error(DH_ERR_REF_UB,
                     "hybridWithPerTriangleFallBack reads row iA < nTriB of A: nTriB > nTriA is "
                     "out of bounds in the reference (hybrid.cpp:130)");
  if (nTriA == 0 || nTriB == 0) return DH_OK;
  static DevBuf geo, meta, pairs_buf, outbuf;
  const size_t nv = 3 * static_cast<size_t>(nTriA + nTriB);
  if ((rc = geo.grow(3 * nv * sizeof(double)))) return rc;
  if ((rc = meta.grow(64 + 2 * sizeof(double) + 9 * sizeof(int)))) return rc;
  if ((rc = pairs_buf.grow(2 * sizeof(int)))) return rc;
  double* gx = geo.as<double>();
  double* gy = gx + nv;
  double* gz = gy + nv;
  const size_t na = 3 * static_cast<size_t>(nTriA), nb = 3 * static_cast<size_t>(nTriB);
  DH_HIP_CHECK(hipMemcpy(gx, xA, na * 8, hipMemcpyHostToDevice));
  DH_HIP_CHECK(hipMemcpy(gx + na, xB, nb * 8, hipMemcpyHostToDevice));
  DH_HIP_CHECK(hipMemcpy(gy, yA, na * 8, hipMemcpyHostToDevice));
  DH_HIP_CHECK(hipMemcpy(gy + na, yB, nb * 8, hipMemcpyHostToDevice));
  DH_HIP_CHECK(hipMemcpy(gz, zA, na * 8, hipMemcpyHostToDevice));
  DH_HIP_CHECK(hipMemcpy(gz + na, zB, nb * 8, hipMemcpyHostToDevice));
  struct Meta {
    double eps[2];
    int off[3];
    int fric[2];
    int id[2];
    int pairs[2];
  } m;
  m.eps[0] = epsA;
  m.eps[1] = epsB;
  m.off[0] = 0;
  m.off[1] = nTriA;
  m.off[2] = nTriA + nTriB;
  m.fric[0] = frictionA ? 1 : 0;
  m.fric[1] = frictionB ? 1 : 0;
  m.id[0] = idA;
  m.id[1] = idB;
  m.pairs[0] = 0;
  m.pairs[1] = 1;
  DH_HIP_CHECK(hipMemcpy(meta.ptr, &m, sizeof(m), hipMemcpyHostToDevice));
  const Meta* dm = meta.as<Meta>();
  dh_geometry g;
  g.x = gx;
  g.y = gy;
  g.z = gz;
  g.tri_offset = dm->off;
  g.eps = dm->eps;
  g.friction = dm->fric;
  g.id = dm->id;
  g.nparticles = 2;
  g.ntri_total = nTriA + nTriB;
  const long long dcap = cap > 0 ? cap : 1;
  if ((rc = outbuf.grow(sizeof(dh_contact) * dcap))) return rc;
  long long total = 0;
  rc = detect_pairs_impl(method, &g, dm->pairs, 1, outbuf.as<dh_contact>(), cap, &total, nullptr,
                         nullptr, stats, nullptr, nullptr, nullptr, nullptr);
  *n_out = static_cast<int>(total);
  if (rc != DH_OK) return rc;
  if (stats && (method == DH_HYBRID_BATCH_STAT || method == DH_HYBRID_TRIANGLE_STAT))
    stats->batch_error = batch_error_sum(stats->batch_error, (epsA + epsB) / 16, nTriA);
  if (total > 0)
    DH_HIP_CHECK(hipMemcpy(out, outbuf.ptr, sizeof(dh_contact) * total, hipMemcpyDeviceToHost));
  return DH_OK;
}

}  // extern "C"
