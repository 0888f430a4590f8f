// pair_features.hpp — single-feature evaluation, optimality certificate and
// canonical frame for the learned narrow phase (host + device, binary64).
//
// A triangle pair's closest points are attained on one of bfSolver's 15
// feature pairs: codes 0..8 are edge k of A against edge m of B (code 3k+m,
// edge k runs from vertex (k+1)%3 towards vertex k, bf.cpp:409-713), codes
// 9..14 are vertex j of B against face A (9+j) and vertex j of A against face
// B (12+j) (bf.cpp:875-1100).  eval_feature reproduces bfSolver's arithmetic
// for one feature, so when the predicted feature is bfSolver's winner the
// closest points are bit-identical to bfSolver's.
//
// certify(): for triangles A, B and candidate points P in A, Q in B with
// n = Q - P, the pair is accepted when n.(a - P) <= tol for every vertex a of
// A and n.(Q - b) <= tol for every vertex b of B, tol = kCertRel * |n|^2.
// Then every x in A and y in B satisfy |y - x| >= |n| - 2 tol / |n|, so
// |n| is within 2*kCertRel of the true distance (and of bfSolver's, which
// equals the true distance for non-intersecting triangles); intersecting
// triangles can never pass.
#pragma once
#include "exact_math.hpp"

namespace dh {

constexpr double kCertRel = 4e-7;  // distance error bound 8e-7 relative

// No feature set: the pair goes to the brute force.
constexpr unsigned kUnresolvedAll = 0xFFFFu;

template <typename T>
DH_HD V3<T> cross3(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Binary32 reciprocal square root (hardware instruction on the device) and
// fused dot product.
DH_HD float rsqrt_frame(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / std::sqrt(x);
#endif
}

DH_HD float fdot(const V3<float>& a, const V3<float>& b) {
  return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}

// Closest points of edge-edge feature `code` (0..8).
DH_HD void eval_ee(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                   V3<double>& Q) {
  {
    const int k = code / 3, m = code - 3 * (code / 3);
    const V3<double> ha = pick3(A, k), oa = pick3(A, k == 2 ? 0 : k + 1);
    const V3<double> hb = pick3(B, m), ob = pick3(B, m == 2 ? 0 : m + 1);
    const V3<double> u = sub(ha, oa), v = sub(hb, ob), w = sub(oa, ob);
    const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    double sc, tc;
    segseg(a, b, c, d, e, sc, tc);
    P = {oa.x + (u.x * sc), oa.y + (u.y * sc), oa.z + (u.z * sc)};
    Q = {ob.x + (v.x * tc), ob.y + (v.y * tc), ob.z + (v.z * tc)};
  }
}

// Closest points of vertex-face feature `code` (9..14).
DH_HD void eval_vf(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                   V3<double>& Q) {
  const int j = code - 9;
  const bool onA = j < 3;
  const V3<double> f0 = sel3(onA, A.v[0], B.v[0]);
  const V3<double> f1 = sel3(onA, A.v[1], B.v[1]);
  const V3<double> f2 = sel3(onA, A.v[2], B.v[2]);
  const V3<double> pt = sel3(onA, pick3(B, j), pick3(A, j - 3));
  const V3<double> e0 = sub(f1, f0), e1 = sub(f2, f0), dd = sub(f0, pt);
  double s, t;
  pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd), dot(dd, dd), s, t);
  const V3<double> onf = {f0.x + ((f2.x - f0.x) * t) + ((f1.x - f0.x) * s),
                          f0.y + ((f2.y - f0.y) * t) + ((f1.y - f0.y) * s),
                          f0.z + ((f2.z - f0.z) * t) + ((f1.z - f0.z) * s)};
  P = sel3(onA, onf, pt);
  Q = sel3(onA, pt, onf);
}

// Closest points of feature `code` (0..14).
DH_HD void eval_feature(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                        V3<double>& Q) {
  if (code < 9)
    eval_ee(A, B, code, P, Q);
  else
    eval_vf(A, B, code, P, Q);
}

// eval_ee, also returning the value bfSolver's edge-edge loop compares for
// this feature: |w + sc u - tc v| (edge_edge, bf.cpp:586-713), and the
// supports of P and Q: the vertex bits of the smallest sub-simplex of A (B)
// that holds the point (a clamped parameter puts it on an edge's end).
DH_HD double eval_ee_value(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                           V3<double>& Q, unsigned& suppA, unsigned& suppB) {
  const int k = code / 3, m = code - 3 * (code / 3);
  const int ko = k == 2 ? 0 : k + 1, mo = m == 2 ? 0 : m + 1;
  const V3<double> ha = pick3(A, k), oa = pick3(A, ko);
  const V3<double> hb = pick3(B, m), ob = pick3(B, mo);
  const V3<double> u = sub(ha, oa), v = sub(hb, ob), w = sub(oa, ob);
  const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
  double sc, tc;
  segseg(a, b, c, d, e, sc, tc);
  P = {oa.x + (u.x * sc), oa.y + (u.y * sc), oa.z + (u.z * sc)};
  Q = {ob.x + (v.x * tc), ob.y + (v.y * tc), ob.z + (v.z * tc)};
  suppA = (sc == 1.0 ? 0u : 1u << ko) | (sc == 0.0 ? 0u : 1u << k);
  suppB = (tc == 1.0 ? 0u : 1u << mo) | (tc == 0.0 ? 0u : 1u << m);
  const V3<double> dP = {w.x + (sc * u.x) - (tc * v.x), w.y + (sc * u.y) - (tc * v.y),
                         w.z + (sc * u.z) - (tc * v.z)};
  return dsqrt(dP.x * dP.x + dP.y * dP.y + dP.z * dP.z);
}

DH_HD double eval_ee_value(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                           V3<double>& Q) {
  unsigned sa, sb;
  return eval_ee_value(A, B, code, P, Q, sa, sb);
}

// Vertex bits (of f0, f1, f2) of the face point pt_tri's parameters (s, t)
// name: P = f0 + t (f2 - f0) + s (f1 - f0); an edge or a vertex when a
// parameter sits on its bound.
DH_HD unsigned face_support(double s, double t) {
  const bool s0 = s == 0.0, t0 = t == 0.0;
  const bool s1 = s == 1.0, t1 = t == 1.0;
  return s1 ? 2u : t1 ? 4u : (s0 && t0) ? 1u : s0 ? 5u : t0 ? 3u : (s + t == 1.0) ? 6u : 7u;
}

// eval_vf, also returning the value bfSolver's vertex-face loop compares:
// the square root of the clamped quadratic form (vertex_face, bf.cpp:875-1100),
// and the supports of P and Q (the vertex alone on its side).
DH_HD double eval_vf_value(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                           V3<double>& Q, unsigned& suppA, unsigned& suppB) {
  const int j = code - 9;
  const bool onA = j < 3;
  const V3<double> f0 = sel3(onA, A.v[0], B.v[0]);
  const V3<double> f1 = sel3(onA, A.v[1], B.v[1]);
  const V3<double> f2 = sel3(onA, A.v[2], B.v[2]);
  const V3<double> pt = sel3(onA, pick3(B, j), pick3(A, j - 3));
  const V3<double> e0 = sub(f1, f0), e1 = sub(f2, f0), dd = sub(f0, pt);
  double s, t;
  const double sqr = pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd),
                            dot(dd, dd), s, t);
  const V3<double> onf = {f0.x + ((f2.x - f0.x) * t) + ((f1.x - f0.x) * s),
                          f0.y + ((f2.y - f0.y) * t) + ((f1.y - f0.y) * s),
                          f0.z + ((f2.z - f0.z) * t) + ((f1.z - f0.z) * s)};
  P = sel3(onA, onf, pt);
  Q = sel3(onA, pt, onf);
  const unsigned fs = face_support(s, t), vs = 1u << (onA ? j : j - 3);
  suppA = onA ? fs : vs;
  suppB = onA ? vs : fs;
  return dsqrt(sqr);
}

DH_HD double eval_vf_value(const Tri<double>& A, const Tri<double>& B, int code, V3<double>& P,
                           V3<double>& Q) {
  unsigned sa, sb;
  return eval_vf_value(A, B, code, P, Q, sa, sb);
}

// ------------------------------------------------------- support choice
//
// support_choice() is the learned path's cheap acceptance test.  P, Q are
// the winner's points and suppA, suppB their supports (the smallest
// sub-simplices holding them).  The pair is accepted when
//   * every vertex outside the supports is lifted off the plane through P
//     (Q) orthogonal to n = Q - P by an angle of at least kLift as seen from
//     P (Q): m.(P - a) >= kLift |a - P|_1 for a in A, m.(b - Q) >= kLift
//     |b - Q|_1 for b in B (m = n / |n|).  The support vertices lie on that
//     plane by construction (the feature's own closest points satisfy its
//     optimality conditions), so with the lifted ones (P, Q) is the unique
//     closest pair of the triangles;
//   * the pair is not touching or intersecting: |n| > 2^-28 (M + S), M the
//     largest coordinate magnitude of P and Q and S the largest vertex
//     distance |a - P|_1, |b - Q|_1 (many zero-valued features tie there);
//   * for an edge-edge winner whose points are both inside their edges, the
//     edges are not parallel: |u x v|^2 >= kLift^2 |u|^2 |v|^2;
//   * neither triangle is degenerate (the same test on two of its edges:
//     a thin triangle's angles are all small or it has two near-parallel
//     edges, so one angle test at vertex 0 rules out coincident and
//     collinear vertices).
// Then the closest pair is unique and isolated, and the features that reach
// it are those whose vertex sets contain both supports (the ties of a
// degenerate support, a vertex against an edge or a vertex): the returned
// set names them, to be evaluated and folded in bfSolver's order.  A feature
// outside the set reaches a value within rounding of the winner's only at
// points within a few 1e-10 of (P, Q) for the lift and the scales of these
// workloads, so bfSolver's choice is reproduced bit for bit unless such a
// near-tie occurs, and its points to that distance always (measured against
// the brute force on every pair; exact_choice() below proves the choice and
// is the exact mode).  Returns kUnresolvedAll when the test fails.
//
// The tests run in binary32 on binary64 differences (vertex - P, Q - P and
// the edges, each rounded once to binary32).  Every binary32 quantity is then
// within a relative 2^-20 of its exact value (a slack m.r within 2^-20 |r|),
// so each threshold is raised by that much: the binary32 test accepts only
// pairs the exact test accepts.
//
// A vertex against a vertex (suppA = {i}, suppB = {j}) ties six features
// (the four edge pairs through both vertices and the two vertex-face pairs).
// With the clamps exact, each of them returns P = A_i and Q = B_j as
// computed by bfSolver: the vertex itself, o + (h - o) for an edge whose far
// end o it starts from, or A_0 + (A_i - A_0) on the face.  When those sums
// reproduce the vertices bit for bit (and no coordinate is zero, whose sign
// the sums could change) every tie gives the same P and Q, and the winner's
// are the answer without evaluating the others.
constexpr float kLift = 1e-3f;

// Features whose vertex sets contain suppA (of A) and suppB (of B).
DH_HD unsigned tie_features(unsigned suppA, unsigned suppB) {
  unsigned out = 0u;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const unsigned ea = (1u << k) | (1u << (k == 2 ? 0 : k + 1));
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const unsigned eb = (1u << m) | (1u << (m == 2 ? 0 : m + 1));
      out |= ((suppA & ~ea) == 0u && (suppB & ~eb) == 0u) ? (1u << (3 * k + m)) : 0u;
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    out |= ((suppB & ~(1u << j)) == 0u) ? (1u << (9 + j)) : 0u;   // vertex j of B, face A
    out |= ((suppA & ~(1u << j)) == 0u) ? (1u << (12 + j)) : 0u;  // vertex j of A, face B
  }
  return out;
}

DH_HD V3<float> to_f32(const V3<double>& v) {
  return {static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)};
}

DH_HD float norm1f(const V3<float>& v) {
  using std::fabs;
  return fabs(v.x) + fabs(v.y) + fabs(v.z);
}

// |u x v|^2 > (kLift^2 + margin) |u|^2 |v|^2 in binary32.
DH_HD bool angle_ok(const V3<float>& u, const V3<float>& v) {
  const V3<float> x = cross3(u, v);
  return fdot(x, x) > (kLift * kLift + 0x1.0p-18f) * (fdot(u, u) * fdot(v, v));
}

// o + (h - o) == h bit for bit, per coordinate, and h has no zero coordinate.
DH_HD bool sum_back(const V3<double>& o, const V3<double>& h) {
  return (o.x + (h.x - o.x)) == h.x && (o.y + (h.y - o.y)) == h.y && (o.z + (h.z - o.z)) == h.z &&
         h.x != 0.0 && h.y != 0.0 && h.z != 0.0;
}

DH_HD unsigned support_choice(const Tri<double>& A, const Tri<double>& B, const V3<double>& P,
                              const V3<double>& Q, unsigned suppA, unsigned suppB, int win) {
  using std::fabs;
  using std::fmax;
  const V3<float> n = to_f32(sub(Q, P));
  const float d2 = fdot(n, n);
  if (!(d2 > 0.f)) return kUnresolvedAll;
  const float inv_d = rsqrt_frame(d2);
  const V3<float> m = {n.x * inv_d, n.y * inv_d, n.z * inv_d};
  const unsigned supp = suppA | (suppB << 3);
  bool ok = true;
  float S = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const V3<float> r = to_f32(k < 3 ? sub(P, A.v[k]) : sub(B.v[k - 3], Q));
    const float r1 = norm1f(r);
    S = fmaxf(S, r1);
    ok = ok && (((supp >> k) & 1u) || fdot(m, r) >= (kLift + 0x1.0p-19f) * r1);
  }
  const float M = fmaxf(fmaxf(fmaxf(fabs(static_cast<float>(P.x)), fabs(static_cast<float>(P.y))),
                              fmaxf(fabs(static_cast<float>(P.z)), fabs(static_cast<float>(Q.x)))),
                        fmaxf(fabs(static_cast<float>(Q.y)), fabs(static_cast<float>(Q.z))));
  // touching or intersecting triangles never pass
  const float lim = 0x1.0p-28f * (M + S);
  ok = ok && d2 > lim * lim;
  // neither triangle degenerate (an angle below kLift at vertex 0)
  ok = ok && angle_ok(to_f32(sub(A.v[1], A.v[0])), to_f32(sub(A.v[2], A.v[0]))) &&
       angle_ok(to_f32(sub(B.v[1], B.v[0])), to_f32(sub(B.v[2], B.v[0])));
  // both points inside their edges: the edges must not be parallel
  const bool interior = win < 9 && __builtin_popcount(suppA) == 2 && __builtin_popcount(suppB) == 2;
  if (interior) {
    const int k = win / 3, mm = win - 3 * (win / 3);
    const V3<float> u = to_f32(sub(pick3(A, k), pick3(A, k == 2 ? 0 : k + 1)));
    const V3<float> v = to_f32(sub(pick3(B, mm), pick3(B, mm == 2 ? 0 : mm + 1)));
    ok = ok && angle_ok(u, v);
  }
  if (!ok) return kUnresolvedAll;
  // a vertex against a vertex whose ties all give the same points
  if (__builtin_popcount(suppA) == 1 && __builtin_popcount(suppB) == 1) {
    const int i = __builtin_ctz(suppA), j = __builtin_ctz(suppB);
    const V3<double> a = pick3(A, i), b = pick3(B, j);
    const bool same = sum_back(pick3(A, i == 2 ? 0 : i + 1), a) && sum_back(A.v[0], a) &&
                      sum_back(pick3(B, j == 2 ? 0 : j + 1), b) && sum_back(B.v[0], b) &&
                      P.x == a.x && P.y == a.y && P.z == a.z && Q.x == b.x && Q.y == b.y &&
                      Q.z == b.z;
    if (same) return 1u << win;
  }
  return tie_features(suppA, suppB);
}


// ---------------------------------------------------------- exact choice
//
// exact_choice() decides, without evaluating the other features, whether the
// feature `win` is the one bfSolver selects.  bfSolver keeps the least value
// among its 9 edge-edge features (first index on ties), the least among its 6
// vertex-face features, and the edge-edge one only when the vertex-face
// minimum is larger (bf.cpp:1102).  `win` and `other` are the two features
// the caller evaluated with bfSolver's own arithmetic (values Vw <= the rule
// against Vo already applied by the caller); P, Q are win's points.  The test
// proves that each of the remaining 13 features has a value above Vw: then
// bfSolver's minima and its final choice are `win`, and its P, Q are win's
// bit for bit.  Returns the set of features (bit g) whose bound does not
// clear Vw: 0 accepts `win` as is; kUnresolvedAll (bit 15 set) means the
// certificate failed (P, Q not a closest pair up to rounding, or P == Q).
// The features in a partial set are the ties of a degenerate closest pair
// (a vertex against an edge or a vertex: several features reach the same
// points, and bfSolver's pick among them follows the last bits of their
// values); resolve_ties() evaluates them.
//
// Lower bounds.  With n = Q - P, d = |n| and the unit normal m = n/d, every
// x in A and y in B satisfy
//     |y - x|^2 = (m.(y - x))^2 + |pi(y - x)|^2,
//     m.(y - x) = d + sA(x) + sB(y),   sA(x) = m.(P - x),  sB(y) = m.(y - Q),
// where pi projects onto the plane orthogonal to m (pi(P) = pi(Q) = 0 after
// translating A by -P and B by -Q).  sA and sB are affine, so over a feature
// (a vertex, an edge or a face: the convex hull of some vertices) they are
// bounded by their vertex values.  The certificate requires sA, sB >= -tol at
// every vertex (then (P, Q) is a closest pair up to rounding).  A vertex is
// "on the plane" when its slack is at most 4 tol, otherwise its slack s > 0
// lifts every point that uses it.  For a feature pair (SA, SB):
//   * if all of SA's vertices are off the plane, sA >= min s over them (same
//     for SB), and |y - x| >= d + those slacks;
//   * otherwise write x = (1-ta) x0 + ta x1 with x0 in the hull of SA's
//     on-plane vertices and x1 in the hull of its off-plane ones: sA(x) >=
//     ta sa (sa = least off-plane slack) and |pi(x) - pi(x0)| <= ta pa (pa =
//     largest in-plane distance from an off-plane to an on-plane vertex).
//     With rho = the in-plane distance between the on-plane hulls of SA and
//     SB, moving the in-plane gap down by u costs at least kappa u of slack,
//     kappa = min(sa/pa, sb/pb), so
//        |y - x|^2 >= min over 0 <= u <= min(rho, U) of
//                     (d + kappa u)^2 + (rho - u)^2,      U = pa + pb,
//     a one-dimensional convex minimum in closed form.
// rho is the distance between points or a point and a segment; a feature
// pair whose on-plane hulls are both segments, or where one is a triangle,
// gets rho = 0 (no in-plane gain claimed), which keeps the bound valid.
//
// Rounding.  The slacks are computed in binary64 and carry an error below
// tol = 2^-48 (L + M + M L / d), L the sum of the largest vertex distances
// from P and from Q, M the largest coordinate magnitude: the vertex
// differences, n = Q - P and the dot products each round at most a few ulp of
// M or of L.  The in-plane geometry (basis, projections, rho, pa) is
// computed in binary32 and widened by 2^-16 L.  bfSolver's computed value of
// a feature is at least its exact distance less 12 eps (L + d) (edge-edge:
// the difference of two points on the segments, the clamped parameters lie
// in [0, 1]) and, for vertex-face, its square at least the exact square less
// 200 eps (L + d)^2 (the quadratic form's terms are at most 20 (L + d)^2).
// A feature is excluded when its bound satisfies
//     LB^2 > (Vw (1 + 2^-40) + 2^-45 (L + Vw))^2 + 2^-42 (L + Vw)^2,
// which dominates both error terms.  NaN inputs fail every comparison and
// are never accepted.
//
// The per-feature bounds are evaluated in binary32 on quantities divided by
// d (so every value is near 1 or a ratio of lengths, far from binary32's
// range limits).  Each input enters as a bound in the safe direction: the
// slacks and dlo rounded down by a factor 1 - 2^-22, the in-plane steps
// rounded up, the threshold rounded up.  The bound is increasing in d, the
// slacks and rho and decreasing in U, so these inputs give a valid lower
// bound.  Its binary32 evaluation (hardware reciprocal and square root, a
// few ulp each, and the in-plane distances' errors below the 2^-16 L
// widening) is then scaled by 1 - 2^-16.  The closed form's branch is
// decided on rounded values; a wrong decision occurs only where the
// unconstrained minimiser u* lies within rounding of 0 or U, and then
// changes the value only at second order, far inside that factor.  A
// feature whose true bound clears the threshold by less than about 2^-15
// stays open and is evaluated, which costs time, never correctness.

struct SideGeom {
  float slack[3];      // slack (m.(P - a) for A, m.(b - Q) for B) less tol, over d, >= 0
  float px[3], py[3];  // in-plane coordinates relative to P (Q), over d
  unsigned on;         // bit k: vertex k on the plane
};

// v[i] for a run-time i in 0..2, as selects (no dynamically indexed arrays,
// which would live in scratch memory on the GPU)
DH_HD float pickf(const float v[3], int i) { return i == 2 ? v[2] : (i == 1 ? v[1] : v[0]); }

// Reciprocal and square root within a few ulp (the hardware's on the GPU).
DH_HD float rcp_ulp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

DH_HD float sqrt_ulp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}

// The coordinate of the vertex named by the single bit `bit` (1, 2 or 4), as
// a sum of selects (the compiler keeps it branch-free; exactly one term is
// nonzero).
DH_HD float pick_bit(const float v[3], unsigned bit) {
  return ((bit & 1u) ? v[0] : 0.f) + ((bit & 2u) ? v[1] : 0.f) + ((bit & 4u) ? v[2] : 0.f);
}

// In-plane distance between the on-plane hulls of vertex sets ma (of A) and
// mb (of B), each one vertex or an edge; 0 when neither is a single vertex.
// Branch-free: the hulls' end points are picked by selects.
DH_HD float inplane_gap(const SideGeom& sa, unsigned ma, const SideGeom& sb, unsigned mb) {
  const unsigned la = ma & (0u - ma), lb = mb & (0u - mb);  // lowest vertex bits
  const unsigned ha = (ma ^ la) ? (ma ^ la) : la;            // the other end (sets of <= 2)
  const unsigned hb = (mb ^ lb) ? (mb ^ lb) : lb;
  // sets are subsets of {1, 2, 4}: single vertices 1, 2, 4; edges 3, 5, 6
  const bool a_single = la != 0u && la == ma, b_single = lb != 0u && lb == mb;
  const bool a_pair = ma == 3u || ma == 5u || ma == 6u;
  const bool b_pair = mb == 3u || mb == 5u || mb == 6u;
  const bool none = !((a_single && (b_single || b_pair)) || (b_single && a_pair));
  const float a0x = pick_bit(sa.px, la), a0y = pick_bit(sa.py, la);
  const float a1x = pick_bit(sa.px, ha), a1y = pick_bit(sa.py, ha);
  const float b0x = pick_bit(sb.px, lb), b0y = pick_bit(sb.py, lb);
  const float b1x = pick_bit(sb.px, hb), b1y = pick_bit(sb.py, hb);
  // p: the single vertex; (q0, q1): the other hull (q0 == q1 for a vertex)
  const float px = a_single ? a0x : b0x, py = a_single ? a0y : b0y;
  const float ax = a_single ? b0x : a0x, ay = a_single ? b0y : a0y;
  const float bx = a_single ? b1x : a1x, by = a_single ? b1y : a1y;
  const float ex = bx - ax, ey = by - ay;
  const float wx = px - ax, wy = py - ay;
  const float ee = ex * ex + ey * ey;
  float t = ee > 0.f ? (wx * ex + wy * ey) * rcp_ulp(ee) : 0.f;
  t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
  const float dx = wx - t * ex, dy = wy - t * ey;
  return none ? 0.f : sqrt_ulp(dx * dx + dy * dy);
}

// Vertex sets of feature `code`: bit k of *ma is vertex k of A (same for B).
DH_HD void feature_vertices(int code, unsigned& ma, unsigned& mb) {
  if (code < 9) {
    const int k = code / 3, m = code - 3 * (code / 3);
    ma = (1u << k) | (1u << (k == 2 ? 0 : k + 1));
    mb = (1u << m) | (1u << (m == 2 ? 0 : m + 1));
  } else if (code < 12) {  // vertex code-9 of B against face A
    ma = 7u;
    mb = 1u << (code - 9);
  } else {  // vertex code-12 of A against face B
    ma = 1u << (code - 12);
    mb = 7u;
  }
}

DH_HD unsigned exact_choice(const Tri<double>& A, const Tri<double>& B, const V3<double>& P,
                            const V3<double>& Q, int win, int other, double Vw) {
  using std::fabs;
  using std::fmax;
  using std::sqrt;
  const V3<double> n = sub(Q, P);
  const double d2 = dot(n, n);
  if (!(d2 > 0.0)) return kUnresolvedAll;
  const double d = sqrt(d2);
  const double inv_d = 1.0 / d;
  const V3<double> m = {n.x * inv_d, n.y * inv_d, n.z * inv_d};
  // in-plane basis (binary32): e1 orthogonal to m (m crossed with the axis
  // least aligned with it), e2 = m x e1
  const float mx = static_cast<float>(m.x), my = static_cast<float>(m.y),
              mz = static_cast<float>(m.z);
  const float ax = fabsf(mx), ay = fabsf(my), az = fabsf(mz);
  const bool xmin = ax <= ay && ax <= az, ymin = !xmin && ay <= az;
  float e1x = xmin ? 0.f : (ymin ? -mz : my);
  float e1y = xmin ? mz : (ymin ? 0.f : -mx);
  float e1z = xmin ? -my : (ymin ? mx : 0.f);
  const float k1 = 1.0f / std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
  e1x *= k1;
  e1y *= k1;
  e1z *= k1;
  const float e2x = my * e1z - mz * e1y, e2y = mz * e1x - mx * e1z, e2z = mx * e1y - my * e1x;
  // per vertex: slack (binary64), in-plane coordinates over d (binary32), sizes
  double La = 0.0, Lb = 0.0;
  double M = fmax(fmax(fmax(fabs(P.x), fabs(P.y)), fmax(fabs(P.z), fabs(Q.x))),
                  fmax(fabs(Q.y), fabs(Q.z)));
  double sla[3], slb[3];
  SideGeom ga, gb;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<double> r = sub(A.v[k], P);
    sla[k] = -dot(m, r);
    La = fmax(La, fabs(r.x) + fabs(r.y) + fabs(r.z));
    M = fmax(M, fmax(fmax(fabs(A.v[k].x), fabs(A.v[k].y)), fabs(A.v[k].z)));
    const float x = static_cast<float>(r.x * inv_d), y = static_cast<float>(r.y * inv_d),
                z = static_cast<float>(r.z * inv_d);
    ga.px[k] = x * e1x + y * e1y + z * e1z;
    ga.py[k] = x * e2x + y * e2y + z * e2z;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<double> r = sub(B.v[k], Q);
    slb[k] = dot(m, r);
    Lb = fmax(Lb, fabs(r.x) + fabs(r.y) + fabs(r.z));
    M = fmax(M, fmax(fmax(fabs(B.v[k].x), fabs(B.v[k].y)), fabs(B.v[k].z)));
    const float x = static_cast<float>(r.x * inv_d), y = static_cast<float>(r.y * inv_d),
                z = static_cast<float>(r.z * inv_d);
    gb.px[k] = x * e1x + y * e1y + z * e1z;
    gb.py[k] = x * e2x + y * e2y + z * e2z;
  }
  const double L = La + Lb;  // the 1-norms bound the 2-norms
  const double tol = 0x1.0p-48 * (L + M + M * L * inv_d);
  ga.on = gb.on = 0u;
  bool cert = true;
  constexpr float kDown = 1.0f - 0x1.0p-22f, kUp = 1.0f + 0x1.0p-22f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    cert = cert && (sla[k] >= -tol) && (slb[k] >= -tol);
    ga.on |= (sla[k] <= 4.0 * tol) ? (1u << k) : 0u;
    gb.on |= (slb[k] <= 4.0 * tol) ? (1u << k) : 0u;
    // less the slacks' rounding, over d, rounded down
    ga.slack[k] = static_cast<float>(fmax(0.0, sla[k] - tol) * inv_d) * kDown;
    gb.slack[k] = static_cast<float>(fmax(0.0, slb[k] - tol) * inv_d) * kDown;
  }
  if (!cert) return kUnresolvedAll;
  // in-plane distances between the vertices of each side (01, 02, 12)
  float da[3], db[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int i = k == 2 ? 1 : 0, j = k == 0 ? 1 : 2;
    const float ux = ga.px[j] - ga.px[i], uy = ga.py[j] - ga.py[i];
    const float vx = gb.px[j] - gb.px[i], vy = gb.py[j] - gb.py[i];
    da[k] = sqrt_ulp(ux * ux + uy * uy);
    db[k] = sqrt_ulp(vx * vx + vy * vy);
  }
  // m.(y - x) >= d - 4 tol when x, y use on-plane vertices (their slacks are
  // >= -tol each, computed within tol), and n's own rounding; over d
  const float dl =
      static_cast<float>(fmax(d * (1.0 - 0x1.0p-50) - 4.0 * tol, 0.0) * inv_d) * kDown;
  const float widen = static_cast<float>(0x1.0p-16 * L * inv_d) * kUp;
  const double lv = L + Vw;
  const double thr_lin = Vw * (1.0 + 0x1.0p-40) + 0x1.0p-45 * lv;
  const float thr =
      static_cast<float>((thr_lin * thr_lin + 0x1.0p-42 * lv * lv) * inv_d * inv_d) * kUp;
  unsigned open = 0u;
#pragma unroll 1
  for (int g = 0; g < 15; ++g) {
    unsigned ma, mb;
    feature_vertices(g, ma, mb);
    const unsigned a0 = ma & ga.on, a1 = ma & ~ga.on, b0 = mb & gb.on, b1 = mb & ~gb.on;
    // least off-plane slack of each side, and the largest in-plane step
    // from an on-plane to an off-plane vertex
    float sa = 3.0e38f, sb = 3.0e38f, pa = 0.f, pb = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const unsigned ak = (a1 >> k) & 1u, bk = (b1 >> k) & 1u;
      sa = ak ? fminf(sa, ga.slack[k]) : sa;
      sb = bk ? fminf(sb, gb.slack[k]) : sb;
      // pair k joins vertices (i, j): an on-plane one and an off-plane one
      const int i = k == 2 ? 1 : 0, j = k == 0 ? 1 : 2;
      const unsigned pa_k = ((a0 >> i) & (a1 >> j) & 1u) | ((a0 >> j) & (a1 >> i) & 1u);
      const unsigned pb_k = ((b0 >> i) & (b1 >> j) & 1u) | ((b0 >> j) & (b1 >> i) & 1u);
      pa = pa_k ? fmaxf(pa, da[k]) : pa;
      pb = pb_k ? fmaxf(pb, db[k]) : pb;
    }
    // a side entirely off the plane: the slab bound
    const float lb = dl + (a0 ? 0.f : sa) + (b0 ? 0.f : sb);
    float lb2 = lb * lb;
    {
      const float rho = fmaxf(0.f, inplane_gap(ga, a0, gb, b0) - widen);
      const float PA = a1 ? (pa + widen) * kUp : 0.f;
      const float PB = b1 ? (pb + widen) * kUp : 0.f;
      // slack per unit of in-plane step (the cheaper side), at most 2^60
      const float ka = PA > 0.f ? sa * rcp_ulp(PA) * kDown : 0x1.0p60f;
      const float kb = PB > 0.f ? sb * rcp_ulp(PB) * kDown : 0x1.0p60f;
      const float kappa = fminf(fminf(ka, kb), 0x1.0p60f);
      const float U = PA + PB;
      // min over 0 <= u <= min(rho, U) of (dl + kappa u)^2 + (rho - u)^2:
      // (dl + kappa rho)^2 / (1 + kappa^2) at u* = (rho - kappa dl) /
      // (1 + kappa^2) when 0 < u* <= U; at u = 0 when u* <= 0; at u = U when
      // u* exceeds U
      const float lead = rho - kappa * dl;
      const float k2 = 1.f + kappa * kappa;
      const bool at0 = U == 0.f || !(lead > 0.f);
      const bool capped = !at0 && U < rho && lead > U * k2;
      const float hU = dl + kappa * U, rU = rho - U, fr = dl + kappa * rho;
      const float lbk = at0 ? dl * dl + rho * rho : (capped ? hU * hU + rU * rU : fr * fr * rcp_ulp(k2));
      lb2 = (a0 && b0) ? lbk : lb2;
    }
    lb2 *= 1.0f - 0x1.0p-16f;
    open |= (g == win || g == other || lb2 > thr) ? 0u : (1u << g);
  }
  return open;
}

DH_HD bool certify(const Tri<double>& A, const Tri<double>& B, const V3<double>& P,
                   const V3<double>& Q) {
  const V3<double> n = sub(Q, P);
  const double d2 = dot(n, n);
  const double tol = kCertRel * d2;
  bool ok = d2 > 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ok = ok && (dot(n, sub(A.v[k], P)) <= tol);
    ok = ok && (dot(n, sub(Q, B.v[k])) <= tol);
  }
  return ok;
}

// The features of `open` besides the two candidates e (edge-edge, value Ve,
// points Pe, Qe) and v (vertex-face) evaluated with bfSolver's arithmetic and
// folded into the two family minima with bfSolver's order (least value,
// lowest code on equal values).  Every feature outside `open` has a value
// above min(Ve, Vv) (exact_choice), so bfSolver's minima are among the
// evaluated ones; the caller then applies bfSolver's final rule
// (edge-edge when Vv > Ve).
DH_HD void fold_feature(const Tri<double>& A, const Tri<double>& B, int g, int& e, double& Ve,
                        V3<double>& Pe, V3<double>& Qe, int& v, double& Vv, V3<double>& Pv,
                        V3<double>& Qv) {
  V3<double> P, Q;
  if (g < 9) {
    const double val = eval_ee_value(A, B, g, P, Q);
    const bool better = val < Ve || (val == Ve && g < e);
    e = better ? g : e;
    Ve = better ? val : Ve;
    Pe = sel3(better, P, Pe);
    Qe = sel3(better, Q, Qe);
  } else {
    const double val = eval_vf_value(A, B, g, P, Q);
    const bool better = val < Vv || (val == Vv && g < v);
    v = better ? g : v;
    Vv = better ? val : Vv;
    Pv = sel3(better, P, Pv);
    Qv = sel3(better, Q, Qv);
  }
}

// (s, t) moved into the parameter triangle s, t >= 0, s + t <= 1.
DH_HD void clamp_bary(double& s, double& t) {
  s = s > 0.0 ? s : 0.0;
  t = t > 0.0 ? t : 0.0;
  const double u = s + t;
  if (u > 1.0) {
    s = s / u;
    t = t / u;
  }
}

// The reference's penalty Newton iteration (the arithmetic of
// penaltySolver(..., int&), penalty.cpp:434-581: regularisation x1e5 from the
// fourth step) started from x0, at most 16 steps, stopping once
// |dx|^2 / |x|^2 < 1e-20 (a Newton residual of 1e-10); the result is moved
// onto the two triangles (parameters clamped into their simplices) so that a
// certificate can bound it.  Returns the number of steps taken.
DH_HD int warm_penalty(const Tri<double>& A, const Tri<double>& B, const double x0[4],
                       V3<double>& P, V3<double>& Q) {
  Penalty<double> pen;
  pen.setup(A, B);
  double x[4] = {x0[0], x0[1], x0[2], x0[3]};
  double reg = pen.reg0;
  int its = 16;
  for (int i = 0; i < 16; ++i) {
    double dx[4];
    reg = i < 3 ? reg : 1E5 * reg;
    pen.step(x, reg, dx);
    x[0] = x[0] - dx[0];
    x[1] = x[1] - dx[1];
    x[2] = x[2] - dx[2];
    x[3] = x[3] - dx[3];
    if (dot4(dx, dx) / dot4(x, x) < 1e-20) {
      its = i + 1;
      break;
    }
  }
  clamp_bary(x[0], x[1]);
  clamp_bary(x[2], x[3]);
  pen.points(x, P, Q);
  return its;
}

// Feature code of bfSolver's winner (same selection as bf_solver).
DH_HD int bf_feature_code(const Tri<double>& A, const Tri<double>& B, V3<double>& P,
                          V3<double>& Q) {
  int ssid = 0, id = 0;
  const bool edge = bf_solver_ids(A, B, P, Q, ssid, id);
  return edge ? ssid : 9 + id;
}

// 1/sqrt(x) for the canonical frame, whose coordinates are rounded to bf16:
// the hardware approximation (v_rsq_f32, about 1 ulp) on the GPU.

// Canonical frame of a pair (origin at A's centroid, x along A's edge 0->1,
// z along A's normal, lengths in units of A's longest edge), as 18 floats:
// A0, A1, A2, B0, B1, B2 (xyz each), clamped to [-32, 32].  The vertices are
// taken relative to A0 in binary64; the frame is built in binary32 with
// fused multiply-adds (the network reads bf16).  Returns false for a
// degenerate A (no frame).
DH_HD bool canonical_inputs(const Tri<double>& A, const Tri<double>& B, float out[18]) {
  V3<float> r[6];
  r[0] = {0.f, 0.f, 0.f};
  for (int v = 1; v < 6; ++v) {
    const V3<double> p = (v < 3) ? A.v[v] : B.v[v - 3];
    r[v] = {static_cast<float>(p.x - A.v[0].x), static_cast<float>(p.y - A.v[0].y),
            static_cast<float>(p.z - A.v[0].z)};
  }
  const V3<float> e01 = r[1], e02 = r[2], e12 = sub(r[2], r[1]);
  const float l01 = fdot(e01, e01), l12 = fdot(e12, e12), l02 = fdot(e02, e02);
  const float lmax = fmaxf(fmaxf(l01, l12), l02);
  const V3<float> nz = cross3(e01, e02);
  const float ln = fdot(nz, nz);
  const bool ok = (l01 > 0.f) && (ln > 0.f);
  const float s = ok ? rsqrt_frame(lmax) : 0.f;
  const float kx = ok ? rsqrt_frame(l01) : 0.f, kz = ok ? rsqrt_frame(ln) : 0.f;
  const V3<float> ex = {e01.x * kx, e01.y * kx, e01.z * kx};
  const V3<float> ez = {nz.x * kz, nz.y * kz, nz.z * kz};
  const V3<float> ey = cross3(ez, ex);
  // scaled axes, and the centroid (r1 + r2) / 3 projected on them
  const V3<float> sx = {ex.x * s, ex.y * s, ex.z * s}, sy = {ey.x * s, ey.y * s, ey.z * s},
                  sz = {ez.x * s, ez.y * s, ez.z * s};
  const V3<float> c = {(r[1].x + r[2].x) * (1.0f / 3.0f), (r[1].y + r[2].y) * (1.0f / 3.0f),
                       (r[1].z + r[2].z) * (1.0f / 3.0f)};
  const float cx = fdot(c, sx), cy = fdot(c, sy), cz = fdot(c, sz);
  for (int v = 0; v < 6; ++v) {
    const float w[3] = {fdot(r[v], sx) - cx, fdot(r[v], sy) - cy, fdot(r[v], sz) - cz};
    for (int a = 0; a < 3; ++a) out[3 * v + a] = fminf(fmaxf(w[a], -32.f), 32.f);
  }
  return ok;
}

// Barycentric parameters of P in A (P = A0 + al (A1-A0) + be (A2-A0)) by the
// 2x2 normal equations.
DH_HD void barycentric(const Tri<double>& A, const V3<double>& P, double& al, double& be) {
  const V3<double> e0 = sub(A.v[1], A.v[0]), e1 = sub(A.v[2], A.v[0]), r = sub(P, A.v[0]);
  const double a = dot(e0, e0), b = dot(e0, e1), c = dot(e1, e1);
  const double d = dot(e0, r), e = dot(e1, r);
  const double det = a * c - b * b;
  const bool ok = det > 0.0;
  al = ok ? (c * d - b * e) / det : 0.0;
  be = ok ? (a * e - b * d) / det : 0.0;
}

}  // namespace dh
