// zoltan_math.hpp — the per-pair arithmetic of the reference's zoltan
// application (zoltan/{bf,penalty,hybrid,contact}.cpp), binary64, in the
// reference's operation order (compile without FMA contraction).  Host and
// device (DH_HD), like exact_math.hpp.
//
//   zoltan_bf          zoltan/bf.cpp:60-952: bfSolver's arithmetic with two early
//                      returns on exactly zero normals and the two switch
//                      fall-throughs (the kZoltan stage of bf_solver_ids).
//   zoltan_penalty     zoltan/penalty.cpp:27-208: 5 Newton steps, eps 0.1,
//                      r = lambda 10^(ln lambda + 14), and a regularisation
//                      multiplied by r before every step once the outer pair
//                      index i reaches 4 (the bug of penalty.cpp:153).
//   zh_penalty         zoltan/hybrid.cpp:134-278: 7 steps, eps 1e-16,
//                      r = lambda 10^(log10 lambda + 8), error output.
//   zh_bf              zoltan/hybrid.cpp:646-885: segment-triangle intersection
//                      first, then 6 point-triangle and 9 segment-segment
//                      distances (SMALL_NUM 1e-8; Q's y and z from the x of the
//                      segment origin, hybrid.cpp:424-426), P from the segments
//                      and Q from the vertices on a tie.
// pow, log and log10 come from the platform's math library: the device's and
// glibc's may differ in the last bit, so the penalty paths agree with the
// reference to rounding, not bit for bit.
#pragma once
#include "exact_math.hpp"

#if !defined(__HIPCC__)
#include <cmath>
#endif

namespace dh {

// zoltan/bf.cpp: returns -1 when the reference returns before writing P and Q
// (normal of B exactly zero, or the mixed test of zoltan/bf.cpp:165-169), else
// 1 when the edge-edge result is kept and 0 otherwise.
DH_HD int zoltan_bf(const Tri<double>& A, const Tri<double>& B, V3<double>& P, V3<double>& Q) {
  {
    const V3<double> u = sub(B.v[1], B.v[0]), v = sub(B.v[2], B.v[1]);
    const double n0 = u.y * v.z - u.z * v.y, n1 = u.z * v.x - u.x * v.z,
                 n2 = u.x * v.y - u.y * v.x;
    if (n0 == 0 && n1 == 0 && n2 == 0) return -1;
    const V3<double> ua = sub(A.v[1], A.v[0]), va = sub(A.v[2], A.v[0]);
    const double m1 = ua.z * va.x - ua.x * va.z, m2 = ua.x * va.y - ua.y * va.x;
    if (n0 == 0 && m1 == 0 && m2 == 0) return -1;
  }
  int ssid, id;
  return bf_solver_ids<double, true>(A, B, P, Q, ssid, id) ? 1 : 0;
}

// Hessian, regularisation seed and the square root of its trace scale.
DH_HD void zoltan_penalty_setup(Penalty<double>& pen, const Tri<double>& A, const Tri<double>& B,
                                double eps, double& delta, double& lambda) {
  pen.setup(A, B);
  const double trace = pen.h[0] + pen.h[5] + pen.h[10] + pen.h[15];
  delta = trace * eps;
  lambda = dsqrt(0.0125 * trace);
}

// zoltan/penalty.cpp for the pair at outer index i.
DH_HD void zoltan_penalty(const Tri<double>& A, const Tri<double>& B, unsigned i, V3<double>& P,
                          V3<double>& Q) {
  Penalty<double> pen;
  double delta, lambda;
  zoltan_penalty_setup(pen, A, B, 0.1, delta, lambda);
  pen.r = lambda * pow(10.0, log(lambda) + 14.0);
  double x[4] = {0.33, 0.33, 0.33, 0.33};
  for (int j = 0; j < 5; ++j) {
    delta = i < 4 ? delta : pen.r * delta;
    double dx[4];
    pen.step(x, delta, dx);
    x[0] = x[0] - dx[0];
    x[1] = x[1] - dx[1];
    x[2] = x[2] - dx[2];
    x[3] = x[3] - dx[3];
  }
  pen.points(x, P, Q);
}

// zoltan/hybrid.cpp:134-278; returns sqrt(|dx|^2 / |x|^2) of the last step.
DH_HD double zh_penalty(const Tri<double>& A, const Tri<double>& B, V3<double>& P,
                        V3<double>& Q) {
  Penalty<double> pen;
  double delta, lambda;
  zoltan_penalty_setup(pen, A, B, 1E-16, delta, lambda);
  pen.r = lambda * pow(10.0, log10(lambda) + 8.0);
  double x[4] = {0.33, 0.33, 0.33, 0.33}, dx[4] = {0, 0, 0, 0};
  for (int i = 0; i < 7; ++i) {
    pen.step(x, delta, dx);  // the 2000x growth of hybrid.cpp:228 needs i >= 7
    x[0] = x[0] - dx[0];
    x[1] = x[1] - dx[1];
    x[2] = x[2] - dx[2];
    x[3] = x[3] - dx[3];
  }
  pen.points(x, P, Q);
  return dsqrt(dot4(dx, dx) / dot4(x, x));
}

// Segment p1-p2 against triangle (a, b, c) (hybrid.cpp:280-342): -1 for a
// degenerate triangle, 2 for a segment in the triangle's plane, 1 when it
// crosses the triangle (P = the crossing), 0 otherwise.  P is written as soon
// as the crossing parameter passes its range test, also when 0 follows.
DH_HD int zh_segt(const V3<double>& p1, const V3<double>& p2, const V3<double>& a,
                  const V3<double>& b, const V3<double>& c, V3<double>& P) {
  const V3<double> u = sub(b, a), v = sub(c, a);
  const V3<double> n = {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
  if (n.x == 0 && n.y == 0 && n.z == 0) return -1;
  const V3<double> dir = sub(p2, p1), w0 = sub(p1, a);
  const double aa = -dot(n, w0);
  const double bb = dot(n, dir);
  if (fabs(bb) < 1E-99) {
    if (aa == 0) return 2;
  }
  const double r = aa / bb;
  if (r < 0.0 || r > 1.0) return 0;  // the r == NAN test of hybrid.cpp:311 is never true
  P = {p1.x + r * dir.x, p1.y + r * dir.y, p1.z + r * dir.z};
  const double uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
  const V3<double> w = sub(P, a);
  const double wu = dot(w, u), wv = dot(w, v);
  const double D = uv * uv - uu * vv;
  const double s = (uv * wv - vv * wu) / D;
  if (s < 0.0 || s > 1.0) return 0;
  const double t = (uv * wu - uu * wv) / D;
  if (t < 0.0 || (s + t) > 1.0) return 0;
  return 1;
}

// Segment p1-p2 against segment p3-p4 (hybrid.cpp:344-445); returns the
// distance of the clamped parameters; P on p2 + (p1 - p2) s, Q = the segment
// p4 + (p3 - p4) t in x, but with p4's x as the origin of y and z.
DH_HD double zh_segseg(const V3<double>& p1, const V3<double>& p2, const V3<double>& p3,
                       const V3<double>& p4, V3<double>& P, V3<double>& Q) {
  const V3<double> u = sub(p1, p2), v = sub(p3, p4), w = sub(p2, p4);
  const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
  const double D = a * c - b * b;
  double sD = D, tD = D, sN = 0, tN = 0;
  const double small = 0.00000001;
  if (D < small) {
    sN = 0.0;
    sD = 1.0;
    tN = e;
    tD = c;
  } else {
    sN = (b * e - c * d);
    tN = (a * e - b * d);
    if (sN < 0.0) {
      sN = 0.0;
      tN = e;
      tD = c;
    } else if (sN > sD) {
      sN = sD;
      tN = e + b;
      tD = c;
    }
  }
  if (tN < 0.0) {
    tN = 0.0;
    if (-d < 0.0) {
      sN = 0.0;
    } else if (-d > a) {
      sN = sD;
    } else {
      sN = -d;
      sD = a;
    }
  } else if (tN > tD) {
    tN = tD;
    if ((-d + b) < 0.0) {
      sN = 0;
    } else if ((-d + b) > a) {
      sN = sD;
    } else {
      sN = (-d + b);
      sD = a;
    }
  }
  const double sc = fabs(sN) < small ? 0.0 : sN / sD;
  const double tc = fabs(tN) < small ? 0.0 : tN / tD;
  P = {p2.x + u.x * sc, p2.y + u.y * sc, p2.z + u.z * sc};
  Q = {p4.x + v.x * tc, p4.x + v.y * tc, p4.x + v.z * tc};
  const V3<double> dP = {w.x + (sc * u.x) - (tc * v.x), w.y + (sc * u.y) - (tc * v.y),
                         w.z + (sc * u.z) - (tc * v.z)};
  return dsqrt(dot(dP, dP));
}

// Point q against triangle (t1, t2, t3) (hybrid.cpp:447-644): the distance;
// tq = the closest point of the triangle.
DH_HD double zh_pt(const V3<double>& t1, const V3<double>& t2, const V3<double>& t3,
                   const V3<double>& q, V3<double>& tq) {
  const V3<double> e0 = sub(t2, t1), e1 = sub(t3, t1), dd = sub(t1, q);
  double s, t;
  const double sqr =
      pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd), dot(dd, dd), s, t);
  tq = {t1.x + (e1.x * t) + (e0.x * s), t1.y + (e1.y * t) + (e0.y * s),
        t1.z + (e1.z * t) + (e0.z * s)};
  return dsqrt(sqr);
}

// zoltan/hybrid.cpp:646-885.  P and Q are in/out: the reference leaves them
// as they came in when no branch of its final selection applies (NaN input).
DH_HD void zh_bf(const V3<double>& A, const V3<double>& B, const V3<double>& C,
                 const V3<double>& D, const V3<double>& E, const V3<double>& F, V3<double>& P,
                 V3<double>& Q) {
  const bool hit = zh_segt(D, E, A, B, C, P) == 1 || zh_segt(E, F, A, B, C, P) == 1 ||
                   zh_segt(F, D, A, B, C, P) == 1 || zh_segt(A, B, D, E, F, P) == 1 ||
                   zh_segt(B, C, D, E, F, P) == 1 || zh_segt(C, A, D, E, F, P) == 1;
  if (hit) {
    Q = P;
    return;
  }
  V3<double> tp, tq, tmp, ssP, ssQ, tmq;
  double ptmin = zh_pt(D, E, F, A, tmp);
  tp = A;
  tq = tmp;
  const V3<double>* vA[3] = {&A, &B, &C};
  const V3<double>* vB[3] = {&D, &E, &F};
  for (int k = 1; k < 3; ++k) {
    const double dd = zh_pt(D, E, F, *vA[k], tmp);
    if (ptmin > dd) {
      ptmin = dd;
      tp = *vA[k];
      tq = tmp;
    }
  }
  for (int k = 0; k < 3; ++k) {
    const double dd = zh_pt(A, B, C, *vB[k], tmp);
    if (ptmin > dd) {
      ptmin = dd;
      tq = *vB[k];
      tp = tmp;
    }
  }
  double ssmin = 1E+30;
  bool first = true;
  for (int ka = 0; ka < 3; ++ka)
    for (int kb = 0; kb < 3; ++kb) {
      const double dd =
          zh_segseg(*vA[ka], *vA[ka == 2 ? 0 : ka + 1], *vB[kb], *vB[kb == 2 ? 0 : kb + 1], tmp, tmq);
      if (first || ssmin > dd) {
        ssmin = dd;
        ssP = tmp;
        ssQ = tmq;
      }
      first = false;
    }
  if (ssmin < ptmin) {
    P = ssP;
    Q = ssQ;
  } else if (ssmin > ptmin) {
    P = tp;
    Q = tq;
  } else if (ssmin == ptmin) {
    P = ssP;
    Q = tq;
  }
}

}  // namespace dh
