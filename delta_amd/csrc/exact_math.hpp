// exact_math.hpp — binary64 triangle-pair kernels for the MI355X engine.
//
// These are the per-pair computations of the reference's bfSolver
// (delta/contact/detection/bf.cpp:217-1178) and penaltySolver
// (penalty.cpp:278-431, 434-581), re-expressed for a SIMT wavefront:
//   * the seven-region point-triangle case analysis is folded into a "kind"
//     code and a single division, so a wave does not diverge over regions;
//   * the nine segment-segment solves are branch-free selects;
//   * every arithmetic operation is the reference's own, in its order, and the
//     translation unit is compiled with -ffp-contract=off, so results are
//     bit-identical to the reference built without FMA.
// The header is also compiled for the host (DH_HD empty) so the same code can
// be checked against the CPU restatement without a GPU.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DH_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define DH_HD inline
#endif

namespace dh {

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
DH_HD V3<T> sub(const V3<T>& a, const V3<T>& b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}

// ((a.x*b.x + a.y*b.y) + a.z*b.z), the reference's DOT macro (algo.h:167)
template <typename T>
DH_HD T dot(const V3<T>& a, const V3<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

template <typename T>
DH_HD T dsqrt(T v) {
  using std::sqrt;
  return sqrt(v);
}

// Clamped segment-segment parameters (bf.cpp:589-663).  Returns sc, tc.
template <typename T>
DH_HD void segseg(T a, T b, T c, T d, T e, T& sc, T& tc) {
  const T small_num = T(1E-30);
  const T D = a * c - b * b;
  const bool par = D < small_num;
  T sN = b * e - c * d;
  T tN = a * e - b * d;
  T sD = D, tD = D;
  // infinite-line solution clamped to the s-range
  const bool s_lo = !par && (sN < T(0));
  const bool s_hi = !par && !(sN < T(0)) && (sN > sD);
  if (par) {
    sN = T(0);
    sD = T(1);
    tN = e;
    tD = c;
  }
  if (s_lo) {
    sN = T(0);
    tN = e;
    tD = c;
  }
  if (s_hi) {
    sN = sD;
    tN = e + b;
    tD = c;
  }
  // clamp t and recompute s on the visible edge (selects, no branches)
  const bool tlo = tN < T(0);
  const bool thi = !tlo && (tN > tD);
  const T v = tlo ? -d : (-d + b);
  const bool lo = v < T(0);
  const bool hi = !lo && (v > a);
  const T nsN = lo ? T(0) : (hi ? sD : v);
  const T nsD = (lo || hi) ? sD : a;
  const bool clamp = tlo || thi;
  sN = clamp ? nsN : sN;
  sD = clamp ? nsD : sD;
  tN = tlo ? T(0) : (thi ? tD : tN);
  using std::fabs;
  const T qs = sN / sD;
  const T qt = tN / tD;
  sc = (fabs(sN) < small_num) ? T(0) : qs;
  tc = (fabs(tN) < small_num) ? T(0) : qt;
}

// Point-triangle squared distance (bf.cpp:922-1091).  The eight outcomes of
// the region analysis: F (s=t=0), S1 (s=1), T1 (t=1), SE (t=0, s=-d/a),
// TE (s=0, t=-e/c), IN (interior), D2 (s+t=1, s=numer/denom), D6 (s+t=1,
// t=numer/denom).

template <typename T>
DH_HD T pt_tri(T a, T b, T c, T d, T e, T f, T& s_out, T& t_out) {
  const T det = a * c - b * b;
  const T s0 = b * e - c * d;
  const T t0 = b * d - a * e;
  // region of (s0, t0) (Eberly's numbering) and the resulting case, as
  // booleans so that the wave selects instead of branching
  const bool c_st = (s0 + t0) <= det;
  const bool sneg = s0 < T(0), tneg = t0 < T(0);
  const bool r0 = c_st && !sneg && !tneg;
  const bool r4 = c_st && sneg && tneg;
  const bool r3 = c_st && sneg && !tneg;
  const bool r5 = c_st && !sneg && tneg;
  const bool r2 = !c_st && sneg;
  const bool r6 = !c_st && !sneg && tneg;
  const bool r1 = !c_st && !sneg && !tneg;
  const T tmp0 = sneg ? (b + d) : (b + e);  // region 2 : region 6
  const T tmp1 = sneg ? (c + e) : (a + d);
  const bool gt = tmp1 > tmp0;
  const T denom = a - T(2) * b + c;
  const T numer26 = tmp1 - tmp0;
  const T numer1 = c + e - b - d;
  const bool nge = numer26 >= denom, t1le = tmp1 <= T(0);
  const bool n1le = numer1 <= T(0), n1ge = numer1 >= denom;
  const bool e_pos = e >= T(0), e_big = -e >= c;
  const bool d_pos = d >= T(0), d_big = -d >= a, d_neg = d < T(0);
  const bool useE = r3 || (r4 && !d_neg) || (r2 && !gt && !t1le);  // s = 0 edge
  const bool useD = r5 || (r6 && !gt && !t1le);                    // t = 0 edge
  const bool bIN = r0;
  const bool bT1 = (useE && !e_pos && e_big) || (r2 && !gt && t1le) || (r6 && gt && nge) ||
                   (r1 && n1le);
  const bool bTE = useE && !e_pos && !e_big;
  const bool bS1 = (useD && !d_pos && d_big) || (r4 && d_neg && d_big) || (r2 && gt && nge) ||
                   (r6 && !gt && t1le) || (r1 && !n1le && n1ge);
  const bool bSE = (useD && !d_pos && !d_big) || (r4 && d_neg && !d_big);
  const bool bD2 = (r2 && gt && !nge) || (r1 && !n1le && !n1ge);
  const bool bD6 = r6 && gt && !nge;
  // one division: 1/det, -d/a, -e/c or numer/denom
  const T num = bIN ? T(1) : bSE ? -d : bTE ? -e : (r1 ? numer1 : numer26);
  const T den = bIN ? det : bSE ? a : bTE ? c : denom;
  const T q = num / den;
  const T s = bIN ? s0 * q : (bD2 || bSE) ? q : bD6 ? T(1) - q : bS1 ? T(1) : T(0);
  const T t = bIN ? t0 * q : bD2 ? T(1) - q : (bD6 || bTE) ? q : bT1 ? T(1) : T(0);
  const T quad = s * (a * s + b * t + T(2) * d) + t * (b * s + c * t + T(2) * e) + f;
  T sqr = (bIN || bD2 || bD6) ? quad
        : bSE ? d * s + f
        : bTE ? e * t + f
        : bS1 ? a + T(2) * d + f
        : bT1 ? c + T(2) * e + f : f;
  sqr = (sqr <= T(0)) ? T(0) : sqr;
  s_out = s;
  t_out = t;
  return sqr;
}

template <typename T>
struct Tri {
  V3<T> v[3];
};

// t.v[i] for a run-time i in 0..2, as selects
template <typename T>
DH_HD V3<T> pick3(const Tri<T>& t, int i) {
  const bool i1 = (i == 1), i2 = (i == 2);
  const T x0 = t.v[0].x, x1 = t.v[1].x, x2 = t.v[2].x;
  const T y0 = t.v[0].y, y1 = t.v[1].y, y2 = t.v[2].y;
  const T z0 = t.v[0].z, z1 = t.v[1].z, z2 = t.v[2].z;
  return {i2 ? x2 : (i1 ? x1 : x0), i2 ? y2 : (i1 ? y1 : y0), i2 ? z2 : (i1 ? z1 : z0)};
}

template <typename T>
DH_HD V3<T> sel3(bool c, const V3<T>& a, const V3<T>& b) {
  return {c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}

// One segment-segment feature of bfSolver (bf.cpp:409-713), folded into the
// running minimum (strict <, first index wins).
template <typename T>
DH_HD void edge_edge(T& ssmin, T& ttc, T& ssc, int& ssid, const V3<T>& ha, const V3<T>& oa,
                     const V3<T>& hb, const V3<T>& ob, int j) {
  const V3<T> u = sub(ha, oa);
  const V3<T> v = sub(hb, ob);
  const V3<T> w = sub(oa, ob);
  const T a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
  T sc, tc;
  segseg(a, b, c, d, e, sc, tc);
  const V3<T> dP = {w.x + (sc * u.x) - (tc * v.x), w.y + (sc * u.y) - (tc * v.y),
                    w.z + (sc * u.z) - (tc * v.z)};
  const T dist = dsqrt(dP.x * dP.x + dP.y * dP.y + dP.z * dP.z);
  const bool better = dist < ssmin;
  ssmin = better ? dist : ssmin;
  ttc = better ? tc : ttc;
  ssc = better ? sc : ssc;
  ssid = better ? j : ssid;
}

// One vertex-face feature (bf.cpp:875-1100): point pt against face f0 f1 f2.
template <typename T>
DH_HD void vertex_face(T& ptmin, T& ss, T& tt, int& id, const V3<T>& f0, const V3<T>& f1,
                       const V3<T>& f2, const V3<T>& pt, int j) {
  const V3<T> e0 = sub(f1, f0);
  const V3<T> e1 = sub(f2, f0);
  const V3<T> dd = sub(f0, pt);
  T s, t;
  const T sqr = pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd),
                       dot(dd, dd), s, t);
  const T dist = dsqrt(sqr);
  const bool better = dist < ptmin;
  ptmin = better ? dist : ptmin;
  ss = better ? s : ss;
  tt = better ? t : tt;
  id = better ? j : id;
}

// bfSolver: closest points P on A and Q on B.  Returns true when the
// edge-edge result is kept (ptmin > ssmin, bf.cpp:1102).  With kZoltan the
// output stage is that of zoltan/bf.cpp, whose switch cases 1 fall through
// into cases 2 (zoltan/bf.cpp:527-559, 907-925): edge-edge feature 1 takes
// its Q from B's edge 2->0 with feature 1's parameter, vertex-face feature 1
// reports B's vertex 2.
template <typename T, bool kZoltan = false>
DH_HD bool bf_solver_ids(const Tri<T>& A, const Tri<T>& B, V3<T>& P, V3<T>& Q, int& ssid_out,
                         int& id_out) {
  T ssmin = T(1E+30), ttc = T(0), ssc = T(0);
  int ssid = 0;
  // Edge k of a triangle runs from its origin v[(k+1)%3] towards v[k].  The
  // features are visited in the reference's order by rotating the vertex
  // triples; the loops are kept rolled so that only one feature is live at a
  // time (register pressure decides the occupancy of this kernel).
  V3<T> a0 = A.v[0], a1 = A.v[1], a2 = A.v[2];
#pragma unroll 1
  for (int k = 0; k < 3; ++k) {
    V3<T> b0 = B.v[0], b1 = B.v[1], b2 = B.v[2];
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      edge_edge(ssmin, ttc, ssc, ssid, a0, a1, b0, b1, 3 * k + m);
      const V3<T> tb = b0;
      b0 = b1;
      b1 = b2;
      b2 = tb;
    }
    const V3<T> ta = a0;
    a0 = a1;
    a1 = a2;
    a2 = ta;
  }
  T ptmin = T(1E30), ss = T(0), tt = T(0);
  int id = 0;
  // j = 0..2: vertex j of B against face A; j = 3..5: vertex j-3 of A against B
  {
    V3<T> p0 = B.v[0], p1 = B.v[1], p2 = B.v[2];
#pragma unroll 1
    for (int j = 0; j < 3; ++j) {
      vertex_face(ptmin, ss, tt, id, A.v[0], A.v[1], A.v[2], p0, j);
      const V3<T> tp = p0;
      p0 = p1;
      p1 = p2;
      p2 = tp;
    }
  }
  {
    V3<T> p0 = A.v[0], p1 = A.v[1], p2 = A.v[2];
#pragma unroll 1
    for (int j = 3; j < 6; ++j) {
      vertex_face(ptmin, ss, tt, id, B.v[0], B.v[1], B.v[2], p0, j);
      const V3<T> tp = p0;
      p0 = p1;
      p1 = p2;
      p2 = tp;
    }
  }
  ssid_out = ssid;
  id_out = id;
  if (ptmin > ssmin) {
    const int k = ssid / 3, m = (kZoltan && ssid == 1) ? 2 : ssid - 3 * (ssid / 3);
    // register selects instead of dynamic indexing (keeps Tri out of scratch)
    const V3<T> ha = pick3(A, k), oa = pick3(A, k == 2 ? 0 : k + 1);
    const V3<T> hb = pick3(B, m), ob = pick3(B, m == 2 ? 0 : m + 1);
    const V3<T> u = sub(ha, oa), v = sub(hb, ob);
    P = {oa.x + (u.x * ssc), oa.y + (u.y * ssc), oa.z + (u.z * ssc)};
    Q = {ob.x + (v.x * ttc), ob.y + (v.y * ttc), ob.z + (v.z * ttc)};
    return true;
  }
  const bool onA = id < 3;
  const V3<T> f0 = sel3(onA, A.v[0], B.v[0]);
  const V3<T> f1 = sel3(onA, A.v[1], B.v[1]);
  const V3<T> f2 = sel3(onA, A.v[2], B.v[2]);
  const V3<T> onf = {f0.x + ((f2.x - f0.x) * tt) + ((f1.x - f0.x) * ss),
                     f0.y + ((f2.y - f0.y) * tt) + ((f1.y - f0.y) * ss),
                     f0.z + ((f2.z - f0.z) * tt) + ((f1.z - f0.z) * ss)};
  const V3<T> vtx = sel3(onA, pick3(B, (kZoltan && id == 1) ? 2 : id), pick3(A, id - 3));
  P = sel3(onA, onf, vtx);
  Q = sel3(onA, vtx, onf);
  return false;
}

template <typename T>
DH_HD bool bf_solver(const Tri<T>& A, const Tri<T>& B, V3<T>& P, V3<T>& Q) {
  int ssid, id;
  return bf_solver_ids(A, B, P, Q, ssid, id);
}

// |Q-P| in the caller's order (bf.cpp:117, contactpoint.cpp:157-159).
template <typename T>
DH_HD T pq_dist(const V3<T>& P, const V3<T>& Q) {
  return dsqrt(((Q.x - P.x) * (Q.x - P.x)) + ((Q.y - P.y) * (Q.y - P.y)) +
               ((Q.z - P.z) * (Q.z - P.z)));
}

// True when one triangle's plane separates the other triangle from it by
// more than the contact margin (no contact, bf.cpp:125): with n = (A1 - A0) x
// (A2 - A0), all of B on one side of A's plane at heights |n.(B_k - A0)| /
// |n| above the margin (or the same with A and B exchanged).  A lies in its
// plane, so the smallest height bounds the distance from below.  Rounding:
// the heights' numerators are within 2^-49.3 |u||v||r| of exact (u, v A's
// edges, r the vertex offsets) and |n| within 2^-49.8 |u||v|; the test
// lowers the heights by 2^-46 |u|_1 |v|_1 |r|_1 and raises |n| by 2^-48
// |u|_1 |v|_1, against the box test's threshold margin (1 + 2^-20) + mag
// 2^-40, beyond which bfSolver's binary64 distance exceeds the margin.
// Two faces of near-touching convex particles are separated by their planes
// almost always when they are apart, so this removes most of the pairs the
// box test leaves at the cost of two cross products.
DH_HD bool plane_side_apart(const Tri<double>& A, const Tri<double>& B, double thr) {
  using std::fabs;
  const V3<double> u = sub(A.v[1], A.v[0]), v = sub(A.v[2], A.v[0]);
  const V3<double> n = {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
  const double uv = (fabs(u.x) + fabs(u.y) + fabs(u.z)) * (fabs(v.x) + fabs(v.y) + fabs(v.z));
  double lo = 1e308, hi = -1e308, R = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<double> r = sub(B.v[k], A.v[0]);
    const double h = dot(n, r);
    lo = fmin(lo, h);
    hi = fmax(hi, h);
    R = fmax(R, fabs(r.x) + fabs(r.y) + fabs(r.z));
  }
  const double E = 0x1p-46 * uv * R;
  const double nn = std::sqrt(dot(n, n)) + 0x1p-48 * uv;
  const double h = lo > 0.0 ? lo : -hi;  // the least height, when all on one side
  return h - E > thr * nn;
}

// The separating axes of the nine edge pairs: n = ea x eb for an edge ea of
// A and eb of B.  Every vertex is projected onto n (relative to A's first
// vertex); when B's projections lie beyond A's by a gap g, every point of B
// is at least g / |n| from every point of A.  n is used as computed (any
// axis bounds the distance), so only the projections round: each is within
// E = 2^-46 |ea|_1 |eb|_1 R of exact (R the largest relative vertex
// distance, as in plane_side_apart), and |n| is bounded above as there.
// True when some axis separates A and B by more than thr.
DH_HD bool edge_axes_apart(const Tri<double>& A, const Tri<double>& B, double thr) {
  using std::fabs;
  V3<double> ra[3], rb[3];
  double R = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ra[k] = sub(A.v[k], A.v[0]);
    rb[k] = sub(B.v[k], A.v[0]);
    R = fmax(R, fmax(fabs(ra[k].x) + fabs(ra[k].y) + fabs(ra[k].z),
                     fabs(rb[k].x) + fabs(rb[k].y) + fabs(rb[k].z)));
  }
  // unrolled: a variable index into A.v or B.v would put both triangles
  // in scratch memory
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<double> ea = sub(A.v[k == 2 ? 0 : k + 1], A.v[k]);
    const double la = fabs(ea.x) + fabs(ea.y) + fabs(ea.z);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const V3<double> eb = sub(B.v[m == 2 ? 0 : m + 1], B.v[m]);
      const V3<double> n = {ea.y * eb.z - ea.z * eb.y, ea.z * eb.x - ea.x * eb.z,
                            ea.x * eb.y - ea.y * eb.x};
      const double uv = la * (fabs(eb.x) + fabs(eb.y) + fabs(eb.z));
      double loA = 1e308, hiA = -1e308, loB = 1e308, hiB = -1e308;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double ha = dot(n, ra[i]), hb = dot(n, rb[i]);
        loA = fmin(loA, ha);
        hiA = fmax(hiA, ha);
        loB = fmin(loB, hb);
        hiB = fmax(hiB, hb);
      }
      const double E = 0x1p-46 * uv * R;
      const double nn = std::sqrt(dot(n, n)) + 0x1p-48 * uv;
      const double gap = fmax(loB - hiA, loA - hiB);
      if (gap - 2.0 * E > thr * nn) return true;
    }
  }
  return false;
}

DH_HD bool plane_apart(const Tri<double>& A, const Tri<double>& B, double margin) {
  using std::fabs;
  double mag = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    mag = fmax(mag, fmax(fmax(fabs(A.v[k].x), fabs(A.v[k].y)),
                         fmax(fmax(fabs(A.v[k].z), fabs(B.v[k].x)),
                              fmax(fabs(B.v[k].y), fabs(B.v[k].z)))));
  const double thr = margin * (1.0 + 0x1p-20) + mag * 0x1p-40;
  return plane_side_apart(A, B, thr) || plane_side_apart(B, A, thr);
}

// plane_apart, then the edge-pair axes (for the pairs the planes leave).
DH_HD bool axes_apart(const Tri<double>& A, const Tri<double>& B, double margin) {
  using std::fabs;
  if (plane_apart(A, B, margin)) return true;
  double mag = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    mag = fmax(mag, fmax(fmax(fabs(A.v[k].x), fabs(A.v[k].y)),
                         fmax(fmax(fabs(A.v[k].z), fabs(B.v[k].x)),
                              fmax(fabs(B.v[k].y), fabs(B.v[k].z)))));
  return edge_axes_apart(A, B, margin * (1.0 + 0x1p-20) + mag * 0x1p-40);
}

// Screen before bfSolver in binary64: true when A and B are provably farther
// apart than the contact margin (no contact, bf.cpp:125).  bfSolver in
// binary32 (coordinates relative to A's first vertex) gives a direction n =
// Q32 - P32; every point of A and B projects onto n, and the gap between the
// two projections divided by |n| bounds the distance from below whatever n
// is.  (The binary32 points themselves may be far from the closest pair: a
// small triangle against a large one, such as a floor, cancels most of the
// bits of the point-face quadratic form, so their distance is no bound.)
// Rounding: the vertices relative to A's first one are rounded once to
// binary32 (within 2^-24 of their length), the projections add 3 ulp, so
// each projection is within 2^-21 |n| S of exact, S the largest relative
// vertex length; the bound is lowered by 2^-19 |n| S and compared, like the
// box test, against margin (1 + 2^-20) + mag 2^-40 (mag the largest
// coordinate magnitude), beyond which bfSolver's binary64 distance exceeds
// the margin as well.
DH_HD bool screened_far(const Tri<double>& A, const Tri<double>& B, double margin) {
  using std::fabs;
  const V3<double> o = A.v[0];
  Tri<float> a32, b32;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    a32.v[k] = {static_cast<float>(A.v[k].x - o.x), static_cast<float>(A.v[k].y - o.y),
                static_cast<float>(A.v[k].z - o.z)};
    b32.v[k] = {static_cast<float>(B.v[k].x - o.x), static_cast<float>(B.v[k].y - o.y),
                static_cast<float>(B.v[k].z - o.z)};
  }
  V3<float> p32, q32;
  bf_solver(a32, b32, p32, q32);
  const V3<float> n = {q32.x - p32.x, q32.y - p32.y, q32.z - p32.z};
  float amax = -3.0e38f, bmin = 3.0e38f, S = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<float> a = a32.v[k], b = b32.v[k];
    amax = fmaxf(amax, n.x * a.x + n.y * a.y + n.z * a.z);
    bmin = fminf(bmin, n.x * b.x + n.y * b.y + n.z * b.z);
    S = fmaxf(S, fmaxf(fabs(a.x) + fabs(a.y) + fabs(a.z), fabs(b.x) + fabs(b.y) + fabs(b.z)));
  }
  const double nn = std::sqrt(static_cast<double>(n.x) * n.x + static_cast<double>(n.y) * n.y +
                              static_cast<double>(n.z) * n.z);
  const double lower = (static_cast<double>(bmin) - static_cast<double>(amax) -
                        0x1p-19 * nn * static_cast<double>(S)) /
                       (nn * (1.0 + 0x1p-20));
  const double mag = fmax(fmax(fabs(o.x), fabs(o.y)), fabs(o.z)) + static_cast<double>(S);
  return nn > 0.0 && lower > margin * (1.0 + 0x1p-20) + mag * 0x1p-40;
}

// Penalty Newton solver state (penalty.cpp:278-357).
template <typename T>
struct Penalty {
  V3<T> A0, B0, BA, CA, ED, FD;
  T h[16];
  T r, reg0;

  DH_HD void setup(const Tri<T>& A, const Tri<T>& B) {
    A0 = A.v[0];
    B0 = B.v[0];
    BA = sub(A.v[1], A.v[0]);
    CA = sub(A.v[2], A.v[0]);
    ED = sub(B.v[1], B.v[0]);
    FD = sub(B.v[2], B.v[0]);
    h[0] = T(2.) * dot(BA, BA);
    h[1] = T(2.) * dot(CA, BA);
    h[2] = T(-2.) * dot(ED, BA);
    h[3] = T(-2.) * dot(FD, BA);
    h[4] = h[1];
    h[5] = T(2.) * dot(CA, CA);
    h[6] = T(-2.) * dot(ED, CA);
    h[7] = T(-2.) * dot(FD, CA);
    h[8] = h[2];
    h[9] = h[6];
    h[10] = T(2.) * dot(ED, ED);
    h[11] = T(2.) * dot(FD, ED);
    h[12] = h[3];
    h[13] = h[7];
    h[14] = h[11];
    h[15] = T(2.) * dot(FD, FD);
    const T trace = h[0] + h[5] + h[10] + h[15];
    reg0 = trace * T(1E-2);
    r = dsqrt(T(0.0125) * trace) * T(1E5);
  }

  // One regularised Newton step from x (penalty.cpp:361-412).
  DH_HD void step(const T x[4], T reg, T dx[4]) const {
    const bool c0 = (-x[0] <= T(0)), c1 = (-x[1] <= T(0)), c2 = (x[0] + x[1] - T(1) <= T(0));
    const bool c3 = (-x[2] <= T(0)), c4 = (-x[3] <= T(0)), c5 = (x[2] + x[3] - T(1) <= T(0));
    T dh[8], mx[6];
    dh[0] = c0 ? T(0) : T(-1);
    mx[0] = c0 ? T(0) : -x[0];
    dh[2] = c1 ? T(0) : T(-1);
    mx[1] = c1 ? T(0) : -x[1];
    dh[1] = dh[3] = c2 ? T(0) : T(1);
    mx[2] = c2 ? T(0) : x[0] + x[1] - T(1);
    dh[4] = c3 ? T(0) : T(-1);
    mx[3] = c3 ? T(0) : -x[2];
    dh[6] = c4 ? T(0) : T(-1);
    mx[4] = c4 ? T(0) : -x[3];
    dh[5] = dh[7] = c5 ? T(0) : T(1);
    mx[5] = c5 ? T(0) : x[2] + x[3] - T(1);
    const V3<T> sxy = {(A0.x + (BA.x * x[0]) + (CA.x * x[1])) - (B0.x + (ED.x * x[2]) + (FD.x * x[3])),
                       (A0.y + (BA.y * x[0]) + (CA.y * x[1])) - (B0.y + (ED.y * x[2]) + (FD.y * x[3])),
                       (A0.z + (BA.z * x[0]) + (CA.z * x[1])) - (B0.z + (ED.z * x[2]) + (FD.z * x[3]))};
    T a[16], b[4];
    b[0] = T(2) * dot(sxy, BA) + r * (dh[0] * mx[0] + dh[1] * mx[2]);
    a[0] = h[0] + r * (dh[0] * dh[0] + dh[1] * dh[1]) + reg;
    a[4] = h[4] + r * (dh[3] * dh[1]);
    const T tmp1 = (h[1] + r * (dh[1] * dh[3])) / a[0];
    a[13] = h[13] - h[12] * tmp1;
    a[9] = h[9] - h[8] * tmp1;
    a[5] = (h[5] + r * (dh[2] * dh[2] + dh[3] * dh[3]) + reg) - a[4] * tmp1;
    b[1] = (T(2) * dot(sxy, CA) + r * (dh[2] * mx[1] + dh[3] * mx[2])) - b[0] * tmp1;
    const T tmp2 = h[2] / a[0];
    const T tmp3 = h[3] / a[0];
    const T tmp4 = ((h[6]) - a[4] * tmp2) / a[5];
    a[14] = ((h[14] + r * (dh[7] * dh[5])) - h[12] * tmp2) - a[13] * tmp4;
    a[10] = ((h[10] + r * (dh[4] * dh[4] + dh[5] * dh[5]) + reg) - h[8] * tmp2) - a[9] * tmp4;
    b[2] = ((T(-2) * dot(sxy, ED) + r * (dh[4] * mx[3] + dh[5] * mx[5])) - b[0] * tmp2) - b[1] * tmp4;
    const T tmp5 = (h[7] - a[4] * tmp3) / a[5];
    const T tmp6 = (((h[11] + r * (dh[5] * dh[7])) - h[8] * tmp3) - a[9] * tmp5) / a[10];
    dx[3] = ((((T(-2) * dot(sxy, FD) + r * (dh[6] * mx[4] + dh[7] * mx[5])) - b[2] * tmp6) -
              b[0] * tmp3) - b[1] * tmp5) /
            ((((h[15] + r * (dh[6] * dh[6] + dh[7] * dh[7]) + reg) - h[12] * tmp3) - a[13] * tmp5) -
             a[14] * tmp6);
    dx[2] = (b[2] - (a[14] * dx[3])) / a[10];
    dx[1] = (b[1] - (a[9] * dx[2] + a[13] * dx[3])) / a[5];
    dx[0] = (b[0] - (a[4] * dx[1] + h[8] * dx[2] + h[12] * dx[3])) / a[0];
  }

  DH_HD void points(const T x[4], V3<T>& P, V3<T>& Q) const {
    P = {A0.x + (BA.x * x[0]) + (CA.x * x[1]), A0.y + (BA.y * x[0]) + (CA.y * x[1]),
         A0.z + (BA.z * x[0]) + (CA.z * x[1])};
    Q = {B0.x + (ED.x * x[2]) + (FD.x * x[3]), B0.y + (ED.y * x[2]) + (FD.y * x[3]),
         B0.z + (ED.z * x[2]) + (FD.z * x[3])};
  }
};

template <typename T>
DH_HD T dot4(const T a[4], const T b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// penaltySolver(..., bool& failed): `iters` fixed steps (4 in the reference),
// regularisation x1e4 after every step, `failed` from the last step.
template <typename T>
DH_HD bool penalty_fixed(const Tri<T>& A, const Tri<T>& B, const T x0[4], int iters, T max_err,
                         V3<T>& P, V3<T>& Q, T xo[4]) {
  Penalty<T> pen;
  pen.setup(A, B);
  T x[4] = {x0[0], x0[1], x0[2], x0[3]};
  T reg = pen.reg0;
  bool failed = false;
  for (int i = 0; i < iters; ++i) {
    T dx[4];
    pen.step(x, reg, dx);
    reg = T(1E4) * reg;
    failed = (dot4(dx, dx) / dot4(x, x) > max_err * max_err);
    x[0] -= dx[0];
    x[1] -= dx[1];
    x[2] -= dx[2];
    x[3] -= dx[3];
  }
  pen.points(x, P, Q);
  if (xo) {
    xo[0] = x[0];
    xo[1] = x[1];
    xo[2] = x[2];
    xo[3] = x[3];
  }
  return failed;
}

// penaltySolver(..., int& its) (penalty.cpp:434-581): <= 16 steps, early exit.
// Returns the iteration count written by the reference, or -1 when it leaves
// the caller's value untouched (no convergence).
template <typename T>
DH_HD int penalty_adaptive(const Tri<T>& A, const Tri<T>& B, T max_err, V3<T>& P, V3<T>& Q) {
  Penalty<T> pen;
  pen.setup(A, B);
  T x[4] = {T(0.33), T(0.33), T(0.33), T(0.33)};
  T reg = pen.reg0;
  int its = -1;
  for (int i = 0; i < 16; ++i) {
    T dx[4];
    reg = i < 3 ? reg : T(1E5) * reg;
    pen.step(x, reg, dx);
    const T err = dot4(dx, dx) / dot4(x, x);
    if (err < max_err * max_err) {
      its = i;
      break;
    }
    x[0] = x[0] - dx[0];
    x[1] = x[1] - dx[1];
    x[2] = x[2] - dx[2];
    x[3] = x[3] - dx[3];
  }
  pen.points(x, P, Q);
  return its;
}

}  // namespace dh
