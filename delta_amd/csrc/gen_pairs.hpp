// gen_pairs.hpp — the seeded workload generator of the isolated narrow-phase
// benchmark, one pair per call (host + device).
//
// Pair i draws from two counter-based splitmix64 streams keyed by (seed, i),
// so every pair is independent of the launch shape.  A: three vertices
// uniform in [0, box)^3.  B, by kind:
//   0 general:      three uniform vertices, centred on A's centroid, then
//                   moved along a uniform random axis by bisection on "not
//                   intersecting and brute-force distance >= s" (the true
//                   distance along t*axis is convex in t and zero at t = 0);
//   1 near-parallel: an edge of B at angle < max_angle to edge A0A1, at
//                   distance s in that edge's outward wedge, B beyond it;
//   2 intersecting: centred on A's centroid (pushed through A).
// s is log-uniform in [sep_min, sep_max].  The device build writes the
// benchmark inputs; the host build (host_check.cpp) lets CPU tests check the
// generator's index ranges and its distribution without a GPU (libm may
// differ from the device's by an ulp, so host pairs are not bit-identical to
// the device's).
#pragma once
#include <cstdint>

#include "exact_math.hpp"
#include "tri_pairs.hpp"

namespace dh {

// splitmix64 (Steele, Lea, Flood 2014): counter-based uniform stream per pair.
DH_HD uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct GenStream {
  uint64_t key;
  uint64_t ctr;
  DH_HD double uniform() {  // [0, 1)
    const uint64_t v = splitmix64(key + (ctr++) * 0x9E3779B97F4A7C15ull);
    return static_cast<double>(v >> 11) * 0x1.0p-53;
  }
};

DH_HD V3<double> gcross(const V3<double>& a, const V3<double>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DH_HD V3<double> gaxpy(double s, const V3<double>& a, const V3<double>& b) {
  return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z};
}
DH_HD V3<double> gscale(double s, const V3<double>& a) { return {s * a.x, s * a.y, s * a.z}; }
DH_HD V3<double> gnormalize(const V3<double>& a) {
  using std::sqrt;
  return gscale(1.0 / sqrt(dot(a, a)), a);
}

// does segment pq cross triangle abc (strictly)?
DH_HD bool seg_crosses_tri(const V3<double>& p, const V3<double>& q, const V3<double>& a,
                           const V3<double>& b, const V3<double>& c) {
  const V3<double> n = gcross(sub(b, a), sub(c, a));
  const double o1 = dot(n, sub(p, a)), o2 = dot(n, sub(q, a));
  if ((o1 > 0 && o2 > 0) || (o1 < 0 && o2 < 0)) return false;
  const V3<double> r = sub(q, p);
  const double v0 = dot(r, gcross(sub(a, p), sub(b, p)));
  const double v1 = dot(r, gcross(sub(b, p), sub(c, p)));
  const double v2 = dot(r, gcross(sub(c, p), sub(a, p)));
  return (v0 >= 0 && v1 >= 0 && v2 >= 0) || (v0 <= 0 && v1 <= 0 && v2 <= 0);
}

DH_HD bool tris_intersect(const Tri<double>& A, const Tri<double>& B) {
  // vertices as named registers (no dynamically indexed arrays)
  const V3<double> a0 = A.v[0], a1 = A.v[1], a2 = A.v[2];
  const V3<double> b0 = B.v[0], b1 = B.v[1], b2 = B.v[2];
  return seg_crosses_tri(a0, a1, b0, b1, b2) || seg_crosses_tri(a1, a2, b0, b1, b2) ||
         seg_crosses_tri(a2, a0, b0, b1, b2) || seg_crosses_tri(b0, b1, a0, a1, a2) ||
         seg_crosses_tri(b1, b2, a0, a1, a2) || seg_crosses_tri(b2, b0, a0, a1, a2);
}

DH_HD double bf_dist(const Tri<double>& A, const Tri<double>& B) {
  V3<double> p, q;
  bf_solver(A, B, p, q);
  return pq_dist(p, q);
}

// Pair i of the workload: triangles A, B, the target separation s and the
// kind (0 general, 1 near-parallel, 2 intersecting).
DH_HD void gen_pair(uint64_t seed, int64_t i, const GenParams& gp, Tri<double>& A,
                    Tri<double>& B, double& s_out, uint8_t& kind_out) {
  using std::cos;
  using std::exp;
  using std::fmax;
  using std::log;
  using std::sin;
  using std::sqrt;
  GenStream ra{splitmix64(seed * 2 + 1) ^ splitmix64(static_cast<uint64_t>(i)), 0};
  GenStream rb{splitmix64(seed * 2 + 2) ^ splitmix64(static_cast<uint64_t>(i) + 0x1234567ull), 0};
  const double L = gp.box;
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
    const V3<double> eh = gscale(1.0 / len, e);
    const V3<double> g = sub(A.v[0], A.v[2]);
    const V3<double> m = gnormalize(gaxpy(-dot(g, eh), eh, g));
    const V3<double> nA = gnormalize(gcross(e, sub(A.v[2], A.v[0])));
    const double phi = (2.0 * rb.uniform() - 1.0) * 1.0;
    const V3<double> w = gaxpy(cos(phi), m, gscale(sin(phi), nA));
    const double tau0 = 0.1 + 0.3 * rb.uniform();
    const V3<double> b0 = gaxpy(s, w, gaxpy(tau0, e, A.v[0]));
    const double th = gp.max_angle * rb.uniform();
    const V3<double> dir = gaxpy(cos(th), eh, gscale(sin(th), gcross(w, eh)));
    const V3<double> b1 = gaxpy((0.3 + 0.3 * rb.uniform()) * len, dir, b0);
    const V3<double> mid = gscale(0.5, V3<double>{b0.x + b1.x, b0.y + b1.y, b0.z + b1.z});
    const V3<double> b2 = gaxpy((0.3 + 0.7 * rb.uniform()) * len, w, mid);
    B.v[0] = b0;
    B.v[1] = b1;
    B.v[2] = b2;
  } else {
    for (int k = 0; k < 3; ++k) B.v[k] = {L * rb.uniform(), L * rb.uniform(), L * rb.uniform()};
    // centre B on A's centroid
    const V3<double> ca = {(A.v[0].x + A.v[1].x + A.v[2].x) / 3,
                           (A.v[0].y + A.v[1].y + A.v[2].y) / 3,
                           (A.v[0].z + A.v[1].z + A.v[2].z) / 3};
    const V3<double> cb = {(B.v[0].x + B.v[1].x + B.v[2].x) / 3,
                           (B.v[0].y + B.v[1].y + B.v[2].y) / 3,
                           (B.v[0].z + B.v[1].z + B.v[2].z) / 3};
    const V3<double> sh = sub(ca, cb);
    for (int k = 0; k < 3; ++k) B.v[k] = gaxpy(1.0, sh, B.v[k]);
    if (u_kind < gp.frac_parallel + gp.frac_intersect) {
      kd = 2;  // pushed through A: centroids coincide
    } else {
      kd = 0;
      const double z = 2.0 * rb.uniform() - 1.0;
      const double ph = 6.283185307179586 * rb.uniform();
      const double rr = sqrt(fmax(0.0, 1.0 - z * z));
      const V3<double> ax = {rr * cos(ph), rr * sin(ph), z};
      double lo = 0.0, hi = s + 4.0 * L;
      for (int it = 0; it < gp.bisection_steps; ++it) {
        const double mid = 0.5 * (lo + hi);
        Tri<double> Bm;
        for (int k = 0; k < 3; ++k) Bm.v[k] = gaxpy(mid, ax, B.v[k]);
        const bool ok = !tris_intersect(A, Bm) && bf_dist(A, Bm) >= s;
        if (ok) hi = mid; else lo = mid;
      }
      for (int k = 0; k < 3; ++k) B.v[k] = gaxpy(hi, ax, B.v[k]);
    }
  }
  s_out = s;
  kind_out = kd;
}

// Index of coordinate c (0..17: A's x0 x1 x2 y0 y1 y2 z0 z1 z2, then B's) of
// pair i in the 18 x n SoA block.
DH_HD int64_t soa_index(int c, int64_t n, int64_t i) { return static_cast<int64_t>(c) * n + i; }

}  // namespace dh
