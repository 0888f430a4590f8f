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

template <typename T>
DH_HD V3<T> cross3(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
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
DH_HD float rsqrt_frame(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / std::sqrt(x);
#endif
}

// Canonical frame of a pair (origin at A's centroid, x along A's edge 0->1,
// z along A's normal, lengths in units of A's longest edge), as 18 floats:
// A0, A1, A2, B0, B1, B2 (xyz each), clamped to [-32, 32].  The vertices are
// taken relative to A0 in binary64; the frame is built in binary32 with
// fused multiply-adds (the network reads bf16).  Returns false for a
// degenerate A (no frame).
DH_HD float fdot(const V3<float>& a, const V3<float>& b) {
  return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}

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
