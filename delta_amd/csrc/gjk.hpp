// gjk.hpp — distance between two convex point sets by the Gilbert–Johnson–
// Keerthi iteration with Johnson's sub-simplex projection (Gilbert, Johnson,
// Keerthi, IEEE J. Robotics and Automation 4(2), 1988), in the arithmetic
// order of the reference's gjk core (delta/contact/detection/gjk.cpp:302-460)
// so that results are bit-identical to it when compiled without FMA.
//
// The support searches are passed in as callables: the host runs them
// sequentially, a wavefront runs them lane-parallel with an (value, index)
// reduction that keeps the first index of the strict extremum, which is the
// sequential scan's choice.  Everything else is uniform across the wave.
#pragma once
#include "exact_math.hpp"

namespace dh {

constexpr double kGjkEpsilon = 1e-6;  // GEOMETRIC_EPSILON, gjk.cpp:102

DH_HD double gdot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Sub-simplices in the order they are tried, as vertex bit sets, and how many
// of them are tried for a simplex of n vertices (gjk.cpp:126-148).
DH_HD int gjk_subset(int m) {
  constexpr int order[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 11, 13, 14, 15};
  return order[m];
}
DH_HD int gjk_subsets_for(int n) { return n == 2 ? 3 : (n == 3 ? 7 : 15); }

// Johnson's projection of the origin onto the simplex w[0..n) (n <= 4).
// On success the simplex is reduced in place to the smallest sub-simplex
// carrying the projection, l[] holds its barycentric weights, v the point,
// and its size is returned; 0 when no sub-simplex qualifies (v, l untouched).
// dl[s][i] holds, for i in s, Johnson's cofactor Delta_i(s), and for j not in
// s, Delta_j(s + {j}).
DH_HD int gjk_project(double w[4][3], int ia[4], int ib[4], int n, double l[4], double v[3]) {
  if (n == 1) {
    l[0] = 1.0;
    v[0] = w[0][0];
    v[1] = w[0][1];
    v[2] = w[0][2];
    return 1;
  }
  double dt[4][4], dl[16][4];
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      dt[i][j] = gdot(w[i], w[j]);
      dt[j][i] = dt[i][j];
    }
  const int tries = gjk_subsets_for(n);
  for (int m = 0; m < tries; ++m) {
    const int s = gjk_subset(m);
    bool ok = true;
    if ((s & (s - 1)) == 0) {  // a single vertex k
      int k = 0;
      while (!(s & (1 << k))) ++k;
      dl[s][k] = 1.0;
      for (int j = 0; j < n; ++j) {
        if (s & (1 << j)) continue;
        dl[s][j] = dt[k][k] - dt[k][j];
        if (dl[s][j] > 0.0) ok = false;
      }
    } else {
      int k = -1;
      for (int i = 0; i < n; ++i) {
        if (!(s & (1 << i))) continue;
        if (k < 0) k = i;
        dl[s][i] = dl[s & ~(1 << i)][i];
        if (dl[s][i] <= 0.0) ok = false;
      }
      for (int j = 0; j < n; ++j) {
        if (s & (1 << j)) continue;
        double acc = 0.0;
        for (int i = 0; i < n; ++i)
          if (s & (1 << i)) acc += dl[s][i] * (dt[i][k] - dt[i][j]);
        dl[s][j] = acc;
        if (acc > 0.0) ok = false;
      }
    }
    if (ok) {
      double sum = 0.0;
      int o = 0;
      for (int i = 0; i < n; ++i) {
        if (!(s & (1 << i))) continue;
        w[o][0] = w[i][0];
        w[o][1] = w[i][1];
        w[o][2] = w[i][2];
        ia[o] = ia[i];
        ib[o] = ib[i];
        l[o] = dl[s][i];
        sum += l[o];
        ++o;
      }
      v[0] = v[1] = v[2] = 0.0;
      for (int i = 0; i < o; ++i) {
        l[i] /= sum;
        v[0] = v[0] + l[i] * w[i][0];
        v[1] = v[1] + l[i] * w[i][1];
        v[2] = v[2] + l[i] * w[i][2];
      }
      return o;
    }
  }
  return 0;
}

// The GJK iteration (gjk.cpp:418-460).  pa(i)/pb(i) return point i of A/B as
// a V3; min_a(v) is the index of the first point of A with the strictly
// smallest dot with v, max_b(v) that of B with the strictly largest (-1 if
// none).  Writes the closest points p in A and q in B and returns the final
// |v|; *ok is false when a support search found no point (NaN input).
template <class PA, class PB, class MinA, class MaxB>
DH_HD double gjk_core(int na, int nb, PA pa, PB pb, MinA min_a, MaxB max_b, double p[3],
                      double q[3], bool* ok) {
  double w[4][3], l[4], v[3];
  int ia[4], ib[4];
  *ok = true;
  const V3<double> a0 = pa(0), b0 = pb(0);
  v[0] = a0.x - b0.x;
  v[1] = a0.y - b0.y;
  v[2] = a0.z - b0.z;
  double vlen = dsqrt(gdot(v, v)), mi = 0.0;
  bool toofar = true;
  int n = 0, iter = 0;
  const long long k = static_cast<long long>(na + nb) * (na + nb);
  while (toofar && vlen > kGjkEpsilon && n < 4 && iter++ < k) {
    const int i = min_a(v), j = max_b(v);
    if (i < 0 || j < 0) {
      *ok = false;
      break;
    }
    const V3<double> A = pa(i), B = pb(j);
    ia[n] = i;
    ib[n] = j;
    w[n][0] = A.x - B.x;
    w[n][1] = A.y - B.y;
    w[n][2] = A.z - B.z;
    const double delta = gdot(v, w[n]) / vlen;
    mi = mi > delta ? mi : delta;
    toofar = (vlen - mi) > kGjkEpsilon;
    if (toofar) {
      n = gjk_project(w, ia, ib, n + 1, l, v);
      vlen = dsqrt(gdot(v, v));
    }
  }
  if (n) {
    p[0] = p[1] = p[2] = 0.0;
    q[0] = q[1] = q[2] = 0.0;
    for (--n; n >= 0; --n) {
      const V3<double> A = pa(ia[n]), B = pb(ib[n]);
      p[0] = p[0] + l[n] * A.x;
      p[1] = p[1] + l[n] * A.y;
      p[2] = p[2] + l[n] * A.z;
      q[0] = q[0] + l[n] * B.x;
      q[1] = q[1] + l[n] * B.y;
      q[2] = q[2] + l[n] * B.z;
    }
  } else {
    p[0] = a0.x;
    p[1] = a0.y;
    p[2] = a0.z;
    q[0] = b0.x;
    q[1] = b0.y;
    q[2] = b0.z;
  }
  return vlen;
}

}  // namespace dh
