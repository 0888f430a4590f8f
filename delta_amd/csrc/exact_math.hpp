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
  // clamp t and recompute s on the visible edge
  const T md = -d;
  const T mdb = (-d + b);
  if (tN < T(0)) {
    tN = T(0);
    const bool lo = md < T(0);
    const bool hi = !lo && (md > a);
    const T nsN = lo ? T(0) : (hi ? sD : md);
    sD = (lo || hi) ? sD : a;
    sN = nsN;
  } else if (tN > tD) {
    tN = tD;
    const bool lo = mdb < T(0);
    const bool hi = !lo && (mdb > a);
    const T nsN = lo ? T(0) : (hi ? sD : mdb);
    sD = (lo || hi) ? sD : a;
    sN = nsN;
  }
  using std::fabs;
  const T qs = sN / sD;
  const T qt = tN / tD;
  sc = (fabs(sN) < small_num) ? T(0) : qs;
  tc = (fabs(tN) < small_num) ? T(0) : qt;
}

// Point-triangle squared distance (bf.cpp:922-1091) with region kinds.
enum PtKind : int { K_F = 0, K_S1, K_T1, K_SE, K_TE, K_IN, K_D2, K_D6 };

template <typename T>
DH_HD T pt_tri(T a, T b, T c, T d, T e, T f, T& s_out, T& t_out) {
  const T det = a * c - b * b;
  const T s0 = b * e - c * d;
  const T t0 = b * d - a * e;
  int kind;
  T numer = T(0), denom = T(1);
  if ((s0 + t0) <= det) {
    if (s0 < T(0)) {
      if (t0 < T(0) && d < T(0)) {
        kind = (-d >= a) ? K_S1 : K_SE;
      } else {
        kind = (e >= T(0)) ? K_F : ((-e >= c) ? K_T1 : K_TE);
      }
    } else if (t0 < T(0)) {
      kind = (d >= T(0)) ? K_F : ((-d >= a) ? K_S1 : K_SE);
    } else {
      kind = K_IN;
    }
  } else {
    if (s0 < T(0)) {
      const T tmp0 = b + d, tmp1 = c + e;
      if (tmp1 > tmp0) {
        numer = tmp1 - tmp0;
        denom = a - T(2) * b + c;
        kind = (numer >= denom) ? K_S1 : K_D2;
      } else {
        kind = (tmp1 <= T(0)) ? K_T1 : ((e >= T(0)) ? K_F : K_TE);
      }
    } else if (t0 < T(0)) {
      const T tmp0 = b + e, tmp1 = a + d;
      if (tmp1 > tmp0) {
        numer = tmp1 - tmp0;
        denom = a - T(2) * b + c;
        kind = (numer >= denom) ? K_T1 : K_D6;
      } else {
        kind = (tmp1 <= T(0)) ? K_S1 : ((d >= T(0)) ? K_F : K_SE);
      }
    } else {
      numer = c + e - b - d;
      if (numer <= T(0)) {
        kind = K_T1;
      } else {
        denom = a - T(2) * b + c;
        kind = (numer >= denom) ? K_S1 : K_D2;
      }
    }
  }
  // one division: 1/det, -d/a, -e/c or numer/denom
  T num = numer, den = denom;
  if (kind == K_IN) {
    num = T(1);
    den = det;
  } else if (kind == K_SE) {
    num = -d;
    den = a;
  } else if (kind == K_TE) {
    num = -e;
    den = c;
  }
  const T q = num / den;
  T s = T(0), t = T(0);
  if (kind == K_IN) {
    s = s0 * q;
    t = t0 * q;
  } else if (kind == K_D2) {
    s = q;
    t = T(1) - q;
  } else if (kind == K_D6) {
    t = q;
    s = T(1) - q;
  } else if (kind == K_SE) {
    s = q;
  } else if (kind == K_TE) {
    t = q;
  } else if (kind == K_S1) {
    s = T(1);
  } else if (kind == K_T1) {
    t = T(1);
  }
  T sqr;
  if (kind == K_IN || kind == K_D2 || kind == K_D6) {
    sqr = s * (a * s + b * t + T(2) * d) + t * (b * s + c * t + T(2) * e) + f;
  } else if (kind == K_SE) {
    sqr = d * s + f;
  } else if (kind == K_TE) {
    sqr = e * t + f;
  } else if (kind == K_S1) {
    sqr = a + T(2) * d + f;
  } else if (kind == K_T1) {
    sqr = c + T(2) * e + f;
  } else {
    sqr = f;
  }
  if (sqr <= T(0)) sqr = T(0);
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
  V3<T> r = t.v[0];
  r = (i == 1) ? t.v[1] : r;
  r = (i == 2) ? t.v[2] : r;
  return r;
}

// bfSolver: closest points P on A and Q on B.  Returns true when the
// edge-edge result is kept (ptmin > ssmin, bf.cpp:1102).
template <typename T>
DH_HD bool bf_solver(const Tri<T>& A, const Tri<T>& B, V3<T>& P, V3<T>& Q) {
  T ssmin = T(1E+30), ttc = T(0), ssc = T(0);
  int ssid = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const V3<T> oa = A.v[(k + 1) % 3];
    const V3<T> u = sub(A.v[k], oa);
    const T a = dot(u, u);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const V3<T> ob = B.v[(m + 1) % 3];
      const V3<T> v = sub(B.v[m], ob);
      const V3<T> w = sub(oa, ob);
      const T b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
      T sc, tc;
      segseg(a, b, c, d, e, sc, tc);
      const V3<T> dP = {w.x + (sc * u.x) - (tc * v.x), w.y + (sc * u.y) - (tc * v.y),
                        w.z + (sc * u.z) - (tc * v.z)};
      const T dist = dsqrt(dP.x * dP.x + dP.y * dP.y + dP.z * dP.z);
      if (dist < ssmin) {
        ssmin = dist;
        ttc = tc;
        ssc = sc;
        ssid = 3 * k + m;
      }
    }
  }
  T ptmin = T(1E30), ss = T(0), tt = T(0);
  int id = 0;
  // j = 0..2: vertex j of B against face A; j = 3..5: vertex j-3 of A against B
  auto vf = [&](const V3<T>& f0, const V3<T>& f1, const V3<T>& f2, const V3<T>& pt, int j) {
    const V3<T> e0 = sub(f1, f0);
    const V3<T> e1 = sub(f2, f0);
    const V3<T> dd = sub(f0, pt);
    T s, t;
    const T sqr = pt_tri(dot(e0, e0), dot(e0, e1), dot(e1, e1), dot(e0, dd), dot(e1, dd),
                         dot(dd, dd), s, t);
    const T dist = dsqrt(sqr);
    if (dist < ptmin) {
      ptmin = dist;
      ss = s;
      tt = t;
      id = j;
    }
  };
  vf(A.v[0], A.v[1], A.v[2], B.v[0], 0);
  vf(A.v[0], A.v[1], A.v[2], B.v[1], 1);
  vf(A.v[0], A.v[1], A.v[2], B.v[2], 2);
  vf(B.v[0], B.v[1], B.v[2], A.v[0], 3);
  vf(B.v[0], B.v[1], B.v[2], A.v[1], 4);
  vf(B.v[0], B.v[1], B.v[2], A.v[2], 5);
  if (ptmin > ssmin) {
    const int k = ssid / 3, m = ssid - 3 * (ssid / 3);
    // register selects instead of dynamic indexing (keeps Tri out of scratch)
    const V3<T> ha = pick3(A, k), oa = pick3(A, k == 2 ? 0 : k + 1);
    const V3<T> hb = pick3(B, m), ob = pick3(B, m == 2 ? 0 : m + 1);
    const V3<T> u = sub(ha, oa), v = sub(hb, ob);
    P = {oa.x + (u.x * ssc), oa.y + (u.y * ssc), oa.z + (u.z * ssc)};
    Q = {ob.x + (v.x * ttc), ob.y + (v.y * ttc), ob.z + (v.z * ttc)};
    return true;
  }
  const bool onA = id < 3;
  const V3<T> f0 = onA ? A.v[0] : B.v[0];
  const V3<T> f1 = onA ? A.v[1] : B.v[1];
  const V3<T> f2 = onA ? A.v[2] : B.v[2];
  const V3<T> onf = {f0.x + ((f2.x - f0.x) * tt) + ((f1.x - f0.x) * ss),
                     f0.y + ((f2.y - f0.y) * tt) + ((f1.y - f0.y) * ss),
                     f0.z + ((f2.z - f0.z) * tt) + ((f1.z - f0.z) * ss)};
  const V3<T> vtx = onA ? pick3(B, id) : pick3(A, id - 3);
  P = onA ? onf : vtx;
  Q = onA ? vtx : onf;
  return false;
}

// |Q-P| in the caller's order (bf.cpp:117, contactpoint.cpp:157-159).
template <typename T>
DH_HD T pq_dist(const V3<T>& P, const V3<T>& Q) {
  return dsqrt(((Q.x - P.x) * (Q.x - P.x)) + ((Q.y - P.y) * (Q.y - P.y)) +
               ((Q.z - P.z) * (Q.z - P.z)));
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
