// ispc_math.hpp — per-pair arithmetic of the reference's ISPC kernels,
// binary64 (REAL = double), in their operation order (compile without FMA
// contraction).  Host and device (DH_HD), like exact_math.hpp.
//
//   ispc_bf        delta/contact/detection/bf.ispc:9-935.  A plane pre-pass
//                  writes a code into distance (2, -2, -1, 0) and the crossing
//                  point into P and Q; 9 segment-segment solves (SMALL_NUM
//                  1e-7) then overwrite P and Q, and the 6 point-triangle
//                  solves overwrite all three only when ptmin < ssmin.  Both
//                  output switches fall through from case 1 into case 2.
//   ispc_penalty   penalty.ispc:3-142: zoltan's 5-step penalty (eps 0.1,
//                  r = lambda 10^(ln lambda + 14)) with the regularisation
//                  multiplied by r before the fifth step only (the Newton index
//                  decides, penalty.ispc:97); returns |P - Q|.
//
// The ISPC sources include an algo.h that is not in the reference; DOT(a, b)
// and SUB(a, b, c) are read as the Solfec-family macros the reference carries
// in zoltan/algo.h:90-96,165-166 (a0*b0 + a1*b1 + a2*b2 in that order; c = a-b).
#pragma once
#include "zoltan_math.hpp"

namespace dh {

// bf.ispc:360-440 for segment p1-p2 (origin p2) against p3-p4 (origin p4).
DH_HD double ispc_segseg(const V3<double>& p1, const V3<double>& p2, const V3<double>& p3,
                         const V3<double>& p4, double& sc, double& tc) {
  const V3<double> u = sub(p1, p2), v = sub(p3, p4), w = sub(p2, p4);
  const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
  const double D = (a * c - b * b);
  double sD = D, tD = D, sN = 0, tN = 0;
  const double small = 0.0000001;
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
  sc = fabs(sN) < small ? 0.0 : sN / sD;
  tc = fabs(tN) < small ? 0.0 : tN / tD;
  const V3<double> dP = {w.x + (sc * u.x) - (tc * v.x), w.y + (sc * u.y) - (tc * v.y),
                         w.z + (sc * u.z) - (tc * v.z)};
  return dsqrt(dP.x * dP.x + dP.y * dP.y + dP.z * dP.z);
}

// bf.ispc:9-935 for one pair, p[0..2] = triangle A (pp1..pp3), p[3..5] = B.
// P, Q and distance are in/out.  A zero normal of B, or the mixed test of
// bf.ispc:110 on the normals' components, writes distance = -1 only.  The
// pre-pass applies DOT to rows of nn[3][2], w0[3][6] and dir[3][6], so for
// every j it reads nn[0][0], nn[0][1], nn[1][0] (j < 3) or nn[1][0],
// nn[1][1], nn[2][0] (j >= 3) and the x components of w0 and dir columns
// 0..2, with A's edges as u and v.  An edge-edge index that is never set
// (every distance NaN or >= 1e30) selects no output case.
DH_HD void ispc_bf(const V3<double> p[6], V3<double>& P, V3<double>& Q, double& distance) {
  // normals: B's (nn[.][0]) and A's (nn[.][1])
  V3<double> u = sub(p[4], p[3]), v = sub(p[5], p[4]);
  const double nb0 = u.y * v.z - u.z * v.y, nb1 = u.z * v.x - u.x * v.z,
               nb2 = u.x * v.y - u.y * v.x;
  if (nb0 == 0 && nb1 == 0 && nb2 == 0) {
    distance = -1;
    return;
  }
  // x components of the first three columns of dir and w0
  const double dir0 = p[1].x - p[0].x, dir1 = p[2].x - p[1].x, dir2 = p[0].x - p[2].x;
  const double w00 = p[0].x - p[3].x, w01 = p[1].x - p[3].x, w02 = p[2].x - p[3].x;
  u = sub(p[1], p[0]);
  v = sub(p[2], p[0]);
  const double na0 = u.y * v.z - u.z * v.y, na1 = u.z * v.x - u.x * v.z;
  const double na2 = u.x * v.y - u.y * v.x;
  if (nb0 == 0 && na1 == 0 && na2 == 0) {
    distance = -1;
    return;
  }
  {
    const double uv = dot(u, v), uu = dot(u, u), vv = dot(v, v);
    for (int j = 0; j < 6; ++j) {
      // DOT(nn[0], .) reads nn[0][0], nn[0][1], nn[1][0]; DOT(nn[1], .) the next three
      const double n0 = j < 3 ? nb0 : nb1, n1 = j < 3 ? na0 : na1, n2 = j < 3 ? nb1 : nb2;
      const double a = -(n0 * w00 + n1 * w01 + n2 * w02);
      const double b = (n0 * dir0 + n1 * dir1 + n2 * dir2);
      if (fabs(b) < 1E-30 && a == 0) {
        distance = 2;
        continue;
      }
      const double r = a / b;
      // column j of dir: the edge from vertex j to the next one of its triangle
      const int nx = j < 3 ? (j + 1) % 3 : 3 + (j - 2) % 3;
      const V3<double> dj = sub(p[nx], p[j]);
      P = {p[j].x + r * dj.x, p[j].y + r * dj.y, p[j].z + r * dj.z};
      Q = P;
      if (r < 0.0 || r > 1.0 || r == INFINITY) {
        distance = -2;
        break;
      }
      const V3<double> w = sub(P, j < 3 ? p[3] : p[0]);
      const double D = uv * uv - uu * vv;
      const double s = (uv * dot(w, v) - vv * dot(w, u)) / D;
      if (s < 0.0 || s > 1.0) {
        distance = -1;
        break;
      }
      const double t = (uv * dot(w, u) - uu * dot(w, v)) / D;
      if (t < 0.0 || (s + t) > 1.0) {
        distance = -1;
        break;
      }
      distance = 0;
    }
  }
  // 9 edge pairs: A's edge k (origin vertex k+1) x B's edge m (origin 3+(m+1)%3)
  double ssmin = 1E+30, ttc = 0, ssc = 0;
  int ssid = -1;
  for (int j = 0; j < 9; ++j) {
    const int ka = j / 3, kb = j - 3 * ka;
    double sc, tc;
    const double dist =
        ispc_segseg(p[ka], p[(ka + 1) % 3], p[3 + kb], p[3 + (kb + 1) % 3], sc, tc);
    if (dist < ssmin) {
      ssmin = dist;
      ttc = tc;
      ssc = sc;
      ssid = j;
    }
  }
  if (ssid >= 0) {
    const int ka = ssid / 3, kb = ssid == 1 ? 2 : ssid - 3 * ka;  // case 1 runs into case 2
    const V3<double> ue = sub(p[ka], p[(ka + 1) % 3]);
    const V3<double> ve = sub(p[3 + kb], p[3 + (kb + 1) % 3]);
    const V3<double>& oa = p[(ka + 1) % 3];
    const V3<double>& ob = p[3 + (kb + 1) % 3];
    P = {oa.x + (ue.x * ssc), oa.y + (ue.y * ssc), oa.z + (ue.z * ssc)};
    Q = {ob.x + (ve.x * ttc), ob.y + (ve.y * ttc), ob.z + (ve.z * ttc)};
  }
  // 6 vertex-face pairs: A's face against B's vertices, then B's against A's
  double ss = 0, tt = 0, ptmin = 1E30;
  int id = 0;
  for (int j = 0; j < 6; ++j) {
    const int f = j < 3 ? 0 : 3;
    const V3<double>& q = p[j < 3 ? 3 + j : j - 3];
    const V3<double> e0 = sub(p[f + 1], p[f]), e1 = sub(p[f + 2], p[f]), dd = sub(p[f], q);
    double s, t;
    const double sqr =
        pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd), dot(dd, dd), s, t);
    const double dist = dsqrt(sqr);
    if (dist < ptmin) {
      ptmin = dist;
      ss = s;
      tt = t;
      id = j;
    }
  }
  if (!(ptmin < ssmin)) return;
  distance = ptmin;
  const int f = id < 3 ? 0 : 3;
  const V3<double> face = {
      p[f].x + ((p[f + 2].x - p[f].x) * tt) + ((p[f + 1].x - p[f].x) * ss),
      p[f].y + ((p[f + 2].y - p[f].y) * tt) + ((p[f + 1].y - p[f].y) * ss),
      p[f].z + ((p[f + 2].z - p[f].z) * tt) + ((p[f + 1].z - p[f].z) * ss)};
  if (id < 3) {
    P = face;
    Q = p[3 + (id == 1 ? 2 : id)];  // case 1 runs into case 2
  } else {
    P = p[id - 3];
    Q = face;
  }
}

// penalty.ispc:3-142.
DH_HD double ispc_penalty(const Tri<double>& A, const Tri<double>& B, V3<double>& P,
                          V3<double>& Q) {
  Penalty<double> pen;
  double delta, lambda;
  zoltan_penalty_setup(pen, A, B, 0.1, delta, lambda);
  pen.r = lambda * pow(10.0, log(lambda) + 14.0);
  double x[4] = {0.33, 0.33, 0.33, 0.33};
  for (int j = 0; j < 5; ++j) {
    delta = j < 4 ? delta : pen.r * delta;
    double dx[4];
    pen.step(x, delta, dx);
    x[0] = x[0] - dx[0];
    x[1] = x[1] - dx[1];
    x[2] = x[2] - dx[2];
    x[3] = x[3] - dx[3];
  }
  pen.points(x, P, Q);
  const V3<double> d = sub(P, Q);
  return dsqrt(dot(d, d));
}

}  // namespace dh
