// host_check.cpp — host build of exact_math.hpp (the GPU kernels' per-pair
// arithmetic) behind plain C entry points, so the kernel math can be compared
// bit-for-bit with the CPU restatement on machines without a GPU.
#include "exact_math.hpp"

using dh::Tri;
using dh::V3;

static Tri<double> load(const double* soa, long n, long i, int base) {
  Tri<double> t;
  for (int k = 0; k < 3; ++k) {
    t.v[k].x = soa[(base + k) * n + i];
    t.v[k].y = soa[(base + 3 + k) * n + i];
    t.v[k].z = soa[(base + 6 + k) * n + i];
  }
  return t;
}

extern "C" void hc_tri_pairs(int method, const double* soa, long n, double eps_total,
                             double* P, double* Q, double* d, unsigned char* flags) {
  const double max_err = eps_total / 16.0;
  const double x0[4] = {0.33, 0.33, 0.33, 0.33};
  for (long i = 0; i < n; ++i) {
    Tri<double> A = load(soa, n, i, 0), B = load(soa, n, i, 9);
    V3<double> p, q;
    unsigned char fl = 0;
    if (method == 0) {
      if (dh::bf_solver(A, B, p, q)) fl |= 4;
    } else if (method == 1 || method == 2) {
      bool f = dh::penalty_fixed(A, B, x0, 4, max_err, p, q, (double*)nullptr);
      if (f) fl |= 1;
      if (method == 2 && f) {
        fl |= 2;
        if (dh::bf_solver(A, B, p, q)) fl |= 4;
      }
    } else {
      dh::penalty_adaptive(A, B, max_err, p, q);
    }
    P[i] = p.x; P[n + i] = p.y; P[2 * n + i] = p.z;
    Q[i] = q.x; Q[n + i] = q.y; Q[2 * n + i] = q.z;
    d[i] = dh::pq_dist(p, q);
    if (flags) flags[i] = fl;
  }
}

#include "dyn_math.hpp"

extern "C" void hc_expmap(double a, double b, double c, double* L) { dh::expmap(a, b, c, L); }
extern "C" void hc_update_rotation(double* w, const double* W, double* R, double h) {
  dh::update_rotation(w, W, R, h);
}
extern "C" void hc_update_angular(double* W, const double* R, const double* J, const double* Ji,
                                  const double* t, double h) {
  dh::update_angular(W, R, J, Ji, t, h);
}
extern "C" void hc_update_vertex(double* x, double* y, double* z, double rx, double ry, double rz,
                                 const double* R, const double* c, const double* C) {
  dh::update_vertex(*x, *y, *z, rx, ry, rz, R, c, C);
}
extern "C" void hc_contact_force(const double* x, const double* n, double depth, int fric,
                                 const double* posA, const double* angA, const double* linA,
                                 double mA, const double* posB, const double* angB,
                                 const double* linB, double mB, double spring, double damper,
                                 double friction, double* force, double* torque) {
  dh::ForceConsts k{spring, damper, friction};
  dh::contact_force(x, n, depth, fric != 0, posA, angA, linA, mA, posB, angB, linB, mB, k, force,
                    torque);
}

#include "pair_features.hpp"

// Per pair: bfSolver's winning feature code, whether eval_feature(code) gives
// bfSolver's points bit for bit, whether that result certifies, and for every
// code 0..14 whether it certifies plus its distance.
extern "C" void hc_learned_probe(const double* soa, long n, int* code, int* same, int* cert_bf,
                                 int* cert_all, double* dist_all, double* d_bf) {
  for (long i = 0; i < n; ++i) {
    Tri<double> A = load(soa, n, i, 0), B = load(soa, n, i, 9);
    V3<double> P, Q, P2, Q2;
    const int c = dh::bf_feature_code(A, B, P, Q);
    code[i] = c;
    dh::eval_feature(A, B, c, P2, Q2);
    same[i] = (P.x == P2.x && P.y == P2.y && P.z == P2.z && Q.x == Q2.x && Q.y == Q2.y &&
               Q.z == Q2.z);
    cert_bf[i] = dh::certify(A, B, P, Q);
    d_bf[i] = dh::pq_dist(P, Q);
    for (int f = 0; f < 15; ++f) {
      dh::eval_feature(A, B, f, P2, Q2);
      cert_all[15 * i + f] = dh::certify(A, B, P2, Q2);
      dist_all[15 * i + f] = dh::pq_dist(P2, Q2);
    }
  }
}

#include "gjk.hpp"

// gjk.hpp's iteration with sequential support searches over AoS point sets
// (the kernel replaces only the searches with a wave reduction).
extern "C" double hc_gjk(const double* a, int na, const double* b, int nb, double* p,
                         double* q) {
  auto pa = [&](int i) { return dh::V3<double>{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; };
  auto pb = [&](int i) { return dh::V3<double>{b[3 * i], b[3 * i + 1], b[3 * i + 2]}; };
  auto min_a = [&](const double* v) {
    int best = -1;
    double bd = 1.7976931348623157e308;
    for (int i = 0; i < na; ++i) {
      const double d = dh::gdot(a + 3 * i, v);
      if (d < bd) {
        bd = d;
        best = i;
      }
    }
    return best;
  };
  auto max_b = [&](const double* v) {
    int best = -1;
    double bd = -1.7976931348623157e308;
    for (int i = 0; i < nb; ++i) {
      const double d = dh::gdot(b + 3 * i, v);
      if (d > bd) {
        bd = d;
        best = i;
      }
    }
    return best;
  };
  bool ok = true;
  const double d = dh::gjk_core(na, nb, pa, pb, min_a, max_b, p, q, &ok);
  return ok ? d : -1.0;
}

#include "zoltan_math.hpp"

// zoltan_math.hpp built for the host, with the kernels' drivers (one pair per
// lane for bf and penalty, one batch of four per lane for hybrid), on
// vertex-major rows D[k*n + i].
extern "C" void hc_zoltan_bf(const double* soa, long n, double* P, double* Q, int* rc) {
  for (long i = 0; i < n; ++i) {
    const Tri<double> A = load(soa, n, i, 0), B = load(soa, n, i, 9);
    dh::V3<double> p, q;
    rc[i] = dh::zoltan_bf(A, B, p, q);
    if (rc[i] >= 0) {
      P[i] = p.x;
      P[n + i] = p.y;
      P[2 * n + i] = p.z;
      Q[i] = q.x;
      Q[n + i] = q.y;
      Q[2 * n + i] = q.z;
    }
  }
}

static Tri<double> hc_tri(const double* abc) {
  Tri<double> t;
  for (int v = 0; v < 3; ++v) t.v[v] = {abc[3 * v], abc[3 * v + 1], abc[3 * v + 2]};
  return t;
}

static Tri<double> hc_row_tri(const double* d, const double* e, const double* f, int n, long i) {
  Tri<double> t;
  t.v[0] = {d[i], d[n + i], d[2 * n + i]};
  t.v[1] = {e[i], e[n + i], e[2 * n + i]};
  t.v[2] = {f[i], f[n + i], f[2 * n + i]};
  return t;
}

static void hc_put(double* p, int n, long i, const dh::V3<double>& v) {
  p[i] = v.x;
  p[n + i] = v.y;
  p[2 * n + i] = v.z;
}

extern "C" void hc_zoltan_penalty(unsigned it, unsigned nt, const double* abc, const double* d,
                                  const double* e, const double* f, double* p, double* q, int n) {
  const Tri<double> A = hc_tri(abc);
  for (unsigned i = it; i < nt; ++i) {
    dh::V3<double> P, Q;
    dh::zoltan_penalty(A, hc_row_tri(d, e, f, n, i), i, P, Q);
    hc_put(p, n, i, P);
    hc_put(q, n, i, Q);
  }
}

extern "C" void hc_zoltan_hybrid(int it, int nt, const double* abc, const double* d,
                                 const double* e, const double* f, double* p, double* q, int n) {
  const Tri<double> A = hc_tri(abc);
  for (unsigned k = static_cast<unsigned>(it); k < static_cast<unsigned>(nt) / 4u; ++k) {
    double err[4];
    Tri<double> B;
    dh::V3<double> P, Q;
    for (unsigned l = 0; l < 4; ++l) {
      B = hc_row_tri(d, e, f, n, k * 4 + l);
      err[l] = dh::zh_penalty(A, B, P, Q);
      hc_put(p, n, k * 4 + l, P);
      hc_put(q, n, k * 4 + l, Q);
    }
    double batch = 0.0;
    for (int l = 0; l < 4; ++l) batch += err[l];
    if (batch / 4u > 1E-8)
      for (unsigned l = 0; l < 4; ++l) {
        dh::zh_bf(A.v[0], A.v[1], A.v[2], B.v[0], B.v[1], B.v[2], P, Q);
        hc_put(p, n, k * 4 + l, P);
        hc_put(q, n, k * 4 + l, Q);
      }
  }
}

#include "ispc_math.hpp"

// ispc_bf / ispc_pen over vertex-major SoA tables pp[v][k][i] (P, Q,
// distance in/out for the brute force, as in the kernel).
extern "C" void hc_ispc(int pen, unsigned n, double* const pp[6][3], double* const P[3],
                        double* const Q[3], double* distance) {
  for (unsigned i = 0; i < n; ++i) {
    V3<double> p[6];
    for (int v = 0; v < 6; ++v) p[v] = {pp[v][0][i], pp[v][1][i], pp[v][2][i]};
    V3<double> Pi = {P[0][i], P[1][i], P[2][i]}, Qi = {Q[0][i], Q[1][i], Q[2][i]};
    if (pen) {
      const Tri<double> A = {{p[0], p[1], p[2]}}, B = {{p[3], p[4], p[5]}};
      distance[i] = dh::ispc_penalty(A, B, Pi, Qi);
    } else {
      dh::ispc_bf(p, Pi, Qi, distance[i]);
    }
    P[0][i] = Pi.x;
    P[1][i] = Pi.y;
    P[2][i] = Pi.z;
    Q[0][i] = Qi.x;
    Q[1][i] = Qi.y;
    Q[2][i] = Qi.z;
  }
}

#include "gen_pairs.hpp"

// Pairs i0 .. i0+m-1 of the unique-pair workload (gen_pairs.hpp) into an
// 18 x m SoA block; returns the number of column indices that fell outside
// [0, 18 m) (always 0: the check the device kernel relies on).
extern "C" long hc_gen_tri_pairs(unsigned long long seed, long i0, long m, double box,
                                 double sep_min, double sep_max, double frac_parallel,
                                 double frac_intersect, double max_angle, int bisection_steps,
                                 double* soa, double* target, unsigned char* kind) {
  dh::GenParams gp{box, sep_min, sep_max, frac_parallel, frac_intersect, max_angle,
                   bisection_steps};
  long bad = 0;
  for (long j = 0; j < m; ++j) {
    Tri<double> A, B;
    double s;
    uint8_t kd;
    dh::gen_pair(seed, i0 + j, gp, A, B, s, kd);
    for (int k = 0; k < 3; ++k) {
      const double v[6] = {A.v[k].x, A.v[k].y, A.v[k].z, B.v[k].x, B.v[k].y, B.v[k].z};
      for (int q = 0; q < 6; ++q) {
        const int64_t at = dh::soa_index(3 * q + k, m, j);
        if (at < 0 || at >= 18 * m) {
          ++bad;
          continue;
        }
        soa[at] = v[q];
      }
    }
    if (target) target[j] = s;
    if (kind) kind[j] = kd;
  }
  return bad;
}

// bfSolver's value of every feature (15 per pair): the numbers its loops
// compare (edge-edge |w + sc u - tc v|, vertex-face sqrt of the quadratic form).
extern "C" void hc_feature_values(const double* soa, long n, double* vals) {
  for (long i = 0; i < n; ++i) {
    Tri<double> A = load(soa, n, i, 0), B = load(soa, n, i, 9);
    V3<double> P, Q;
    for (int f = 0; f < 15; ++f)
      vals[15 * i + f] = f < 9 ? dh::eval_ee_value(A, B, f, P, Q) : dh::eval_vf_value(A, B, f, P, Q);
  }
}

// The learned path's decision for given candidates (the network's most likely
// edge-edge feature ce[i] and vertex-face feature cv[i]), as k_learned makes
// it: both evaluated, bfSolver's rule, the acceptance test, ties folded.  status:
// 0 sent to the brute force (P, Q untouched), 1 accepted as is, 2 + k
// accepted after folding k tied features; P, Q (component-major) the result.
// on: on-plane vertex counts at the result (A in bits 0-1, B in bits 2-3).
static void learned_decide(int kind, const double* soa, long n, const int* ce, const int* cv,
                           double* Pout, double* Qout, int* status) {
  const bool exact = kind == 1;
  for (long i = 0; i < n; ++i) {
    Tri<double> A = load(soa, n, i, 0), B = load(soa, n, i, 9);
    int fe = ce[i], fv = cv[i];
    V3<double> Pe, Qe, Pv, Qv;
    unsigned sAe, sBe, sAv, sBv;
    double Ve = dh::eval_ee_value(A, B, fe, Pe, Qe, sAe, sBe);
    double Vv = dh::eval_vf_value(A, B, fv, Pv, Qv, sAv, sBv);
    const bool take_e = Vv > Ve;
    const int win0 = take_e ? fe : fv;
    V3<double> P = take_e ? Pe : Pv, Q = take_e ? Qe : Qv;
    const unsigned open =
        exact ? dh::exact_choice(A, B, P, Q, win0, take_e ? fv : fe, take_e ? Ve : Vv)
              : dh::support_choice(A, B, P, Q, take_e ? sAe : sAv, take_e ? sBe : sBv, win0);
    if (open == dh::kUnresolvedAll) {
      status[i] = 0;
      continue;
    }
    const unsigned rest = open & ~((1u << fe) | (1u << fv));
    int k = 0;
    for (int g = 0; g < 15; ++g)
      if ((rest >> g) & 1u) {
        dh::fold_feature(A, B, g, fe, Ve, Pe, Qe, fv, Vv, Pv, Qv);
        ++k;
      }
    const bool te = Vv > Ve;
    P = te ? Pe : Pv;
    Q = te ? Qe : Qv;
    status[i] = rest ? 2 + k : 1;
    Pout[i] = P.x; Pout[n + i] = P.y; Pout[2 * n + i] = P.z;
    Qout[i] = Q.x; Qout[n + i] = Q.y; Qout[2 * n + i] = Q.z;
  }
}

// exact_choice (the proof of bfSolver's choice) and support_choice (the
// learned path's test).
extern "C" void hc_learned_decide(const double* soa, long n, const int* ce, const int* cv,
                                  double* Pout, double* Qout, int* status) {
  learned_decide(1, soa, n, ce, cv, Pout, Qout, status);
}

extern "C" void hc_learned_decide_support(const double* soa, long n, const int* ce,
                                          const int* cv, double* Pout, double* Qout,
                                          int* status) {
  learned_decide(0, soa, n, ce, cv, Pout, Qout, status);
}


// screened_far (the binary32 screen before the binary64 bfSolver) per pair.
extern "C" void hc_screened_far(const double* soa, long n, double margin, unsigned char* far) {
  for (long i = 0; i < n; ++i) far[i] = dh::screened_far(load(soa, n, i, 0), load(soa, n, i, 9), margin) ? 1 : 0;
}

// axes_apart (the triangle planes, then the edge-pair axes) per pair.
extern "C" void hc_axes_apart(const double* soa, long n, double margin, unsigned char* far) {
  for (long i = 0; i < n; ++i) far[i] = dh::axes_apart(load(soa, n, i, 0), load(soa, n, i, 9), margin) ? 1 : 0;
}

// plane_apart (the triangle-plane test before the binary64 bfSolver) per pair.
extern "C" void hc_plane_apart(const double* soa, long n, double margin, unsigned char* far) {
  for (long i = 0; i < n; ++i) far[i] = dh::plane_apart(load(soa, n, i, 0), load(soa, n, i, 9), margin) ? 1 : 0;
}

// cr_sincos (the device's sin and cos in expmap's large-angle branch), host build.
extern "C" void hc_cr_sincos(const double* x, long n, double* s, double* c) {
  for (long i = 0; i < n; ++i) {
    s[i] = dh::cr_sincos(x[i], false);
    c[i] = dh::cr_sincos(x[i], true);
  }
}
