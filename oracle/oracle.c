/*
 * oracle.c — CPU restatement of the reference path (test infrastructure only;
 * see oracle.h for what may call it and how each function is pinned).
 *
 * Compile with -O2 -ffp-contract=off: the reference's gcc build (delta/Makefile
 * :22-23,41) has no -march and hence no fused multiply-add, so every product
 * below is rounded before it is added, in the order written.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_DOT(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2])

static void sub3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}

/* ---------------------------------------------------------------------------
 * Segment-segment part of bfSolver (bf.cpp:409-713).  The nine edge pairs are
 * enumerated as (edge k of A) x (edge m of B), j = 3k+m.  Edge k of a triangle
 * with vertices v0,v1,v2 runs from its origin v[(k+1)%3] towards v[k]: the
 * reference's u = p1-p2 (origin p2), p2-p3 (origin p3), p3-p1 (origin p1),
 * and likewise v = p4-p5 (origin p5), p5-p6, p6-p4.
 * The clamped solve is the classic segment-segment closest-point routine with
 * the reference's thresholds (D < 1e-30, |sN|,|tN| < 1e-30 -> 0).
 * ------------------------------------------------------------------------- */
static void segseg_params(double a, double b, double c, double d, double e,
                          double* sc_out, double* tc_out) {
  const double small_num = 1E-30;
  double D = a * c - b * b;
  double sD = D, tD = D, sN, tN;
  if (D < small_num) {
    sN = 0.0;
    sD = 1.0;
    tN = e;
    tD = c;
  } else {
    sN = b * e - c * d;
    tN = a * e - b * d;
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
  *sc_out = (fabs(sN) < small_num) ? 0.0 : sN / sD;
  *tc_out = (fabs(tN) < small_num) ? 0.0 : tN / tD;
}

/* ---------------------------------------------------------------------------
 * Point-triangle part (bf.cpp:875-1100): squared distance from a point to the
 * triangle base + s*E0 + t*E1 with the seven-region case analysis, given
 * a=E0.E0, b=E0.E1, c=E1.E1, d=E0.(base-pt), e=E1.(base-pt), f=|base-pt|^2.
 * ------------------------------------------------------------------------- */
static double pt_tri_sqr(double a, double b, double c, double d, double e, double f,
                         double* s_out, double* t_out) {
  double det = a * c - b * b;
  double s = b * e - c * d;
  double t = b * d - a * e;
  double sqr = 0.0;
#define QUAD_FORM (s * (a * s + b * t + 2 * d) + t * (b * s + c * t + 2 * e) + f)
  if ((s + t) <= det) {
    if (s < 0) {
      if (t < 0) {
        if (d < 0) {
          t = 0;
          if (-d >= a) {
            s = 1;
            sqr = a + 2 * d + f;
          } else {
            s = -d / a;
            sqr = d * s + f;
          }
        } else {
          s = 0;
          if (e >= 0) {
            t = 0;
            sqr = f;
          } else if (-e >= c) {
            t = 1;
            sqr = c + 2 * e + f;
          } else {
            t = -e / c;
            sqr = e * t + f;
          }
        }
      } else {
        s = 0;
        if (e >= 0) {
          t = 0;
          sqr = f;
        } else if (-e >= c) {
          t = 1;
          sqr = c + 2 * e + f;
        } else {
          t = -e / c;
          sqr = e * t + f;
        }
      }
    } else if (t < 0) {
      t = 0;
      if (d >= 0) {
        s = 0;
        sqr = f;
      } else if (-d >= a) {
        s = 1;
        sqr = a + 2 * d + f;
      } else {
        s = -d / a;
        sqr = d * s + f;
      }
    } else {
      double inv_det = 1 / det;
      s = s * inv_det;
      t = t * inv_det;
      sqr = QUAD_FORM;
    }
  } else {
    if (s < 0) {
      double tmp0 = b + d;
      double tmp1 = c + e;
      if (tmp1 > tmp0) {
        double numer = tmp1 - tmp0;
        double denom = a - 2 * b + c;
        if (numer >= denom) {
          s = 1;
          t = 0;
          sqr = a + 2 * d + f;
        } else {
          s = numer / denom;
          t = 1 - s;
          sqr = QUAD_FORM;
        }
      } else {
        s = 0;
        if (tmp1 <= 0) {
          t = 1;
          sqr = c + 2 * e + f;
        } else if (e >= 0) {
          t = 0;
          sqr = f;
        } else {
          t = -e / c;
          sqr = e * t + f;
        }
      }
    } else if (t < 0) {
      double tmp0 = b + e;
      double tmp1 = a + d;
      if (tmp1 > tmp0) {
        double numer = tmp1 - tmp0;
        double denom = a - 2 * b + c;
        if (numer >= denom) {
          t = 1;
          s = 0;
          sqr = c + 2 * e + f;
        } else {
          t = numer / denom;
          s = 1 - t;
          sqr = QUAD_FORM;
        }
      } else {
        t = 0;
        if (tmp1 <= 0) {
          s = 1;
          sqr = a + 2 * d + f;
        } else if (d >= 0) {
          s = 0;
          sqr = f;
        } else {
          s = -d / a;
          sqr = d * s + f;
        }
      }
    } else {
      double numer = c + e - b - d;
      if (numer <= 0) {
        s = 0;
        t = 1;
        sqr = c + 2 * e + f;
      } else {
        double denom = a - 2 * b + c;
        if (numer >= denom) {
          s = 1;
          t = 0;
          sqr = a + 2 * d + f;
        } else {
          s = numer / denom;
          t = 1 - s;
          sqr = QUAD_FORM;
        }
      }
    }
  }
#undef QUAD_FORM
  if (sqr <= 0.0) sqr = 0.0;
  *s_out = s;
  *t_out = t;
  return sqr;
}

/* bfSolver on vertex tables; `zoltan` selects the output stage of
 * zoltan/bf.cpp, whose cases 1 fall through into cases 2 (527-559, 907-925). */
static int bf_core(const double va[3][3], const double vb[3][3], double P[3], double Q[3],
                   int zoltan);

int or_bf_solver(const double* xA, const double* yA, const double* zA,
                 const double* xB, const double* yB, const double* zB,
                 double P[3], double Q[3]) {
  double va[3][3], vb[3][3];
  for (int k = 0; k < 3; ++k) {
    va[k][0] = xA[k]; va[k][1] = yA[k]; va[k][2] = zA[k];
    vb[k][0] = xB[k]; vb[k][1] = yB[k]; vb[k][2] = zB[k];
  }
  return bf_core(va, vb, P, Q, 0);
}

/* zoltan/bf.cpp:60-952: -1 when it returns before writing P, Q (exactly zero
 * normal of B, or the mixed test nB.x == nA.y == nA.z == 0 of 165-169). */
int or_zoltan_bf(const double* xA, const double* yA, const double* zA,
                 const double* xB, const double* yB, const double* zB,
                 double P[3], double Q[3]) {
  double va[3][3], vb[3][3];
  for (int k = 0; k < 3; ++k) {
    va[k][0] = xA[k]; va[k][1] = yA[k]; va[k][2] = zA[k];
    vb[k][0] = xB[k]; vb[k][1] = yB[k]; vb[k][2] = zB[k];
  }
  double u[3], v[3];
  sub3(vb[1], vb[0], u);
  sub3(vb[2], vb[1], v);
  const double nb0 = u[1] * v[2] - u[2] * v[1], nb1 = u[2] * v[0] - u[0] * v[2],
               nb2 = u[0] * v[1] - u[1] * v[0];
  if (nb0 == 0 && nb1 == 0 && nb2 == 0) return -1;
  sub3(va[1], va[0], u);
  sub3(va[2], va[0], v);
  const double na1 = u[2] * v[0] - u[0] * v[2], na2 = u[0] * v[1] - u[1] * v[0];
  if (nb0 == 0 && na1 == 0 && na2 == 0) return -1;
  return bf_core(va, vb, P, Q, 1);
}

static int bf_core(const double va[3][3], const double vb[3][3], double P[3], double Q[3],
                   int zoltan) {
  /* The plane/segment pre-pass at bf.cpp:270-407 writes the outputs and is
   * always overwritten below; it has no observable effect and is omitted. */

  /* edge-edge */
  double ssmin = 1E+30, ttc = 0, ssc = 0;
  int ssid = 0;
  for (int k = 0; k < 3; ++k) {
    double u[3];
    sub3(va[k], va[(k + 1) % 3], u);
    const double a = OR_DOT(u, u);
    for (int m = 0; m < 3; ++m) {
      double v[3], w[3];
      sub3(vb[m], vb[(m + 1) % 3], v);
      sub3(va[(k + 1) % 3], vb[(m + 1) % 3], w);
      double b = OR_DOT(u, v), c = OR_DOT(v, v), d = OR_DOT(u, w), e = OR_DOT(v, w);
      double sc, tc;
      segseg_params(a, b, c, d, e, &sc, &tc);
      double dP[3];
      for (int q = 0; q < 3; ++q) dP[q] = w[q] + (sc * u[q]) - (tc * v[q]);
      double dist = sqrt(dP[0] * dP[0] + dP[1] * dP[1] + dP[2] * dP[2]);
      if (dist < ssmin) {
        ssmin = dist;
        ttc = tc;
        ssc = sc;
        ssid = 3 * k + m;
      }
    }
  }
  {
    int k = ssid / 3, m = (zoltan && ssid == 1) ? 2 : ssid % 3;
    const double* oa = va[(k + 1) % 3];
    const double* ob = vb[(m + 1) % 3];
    double u[3], v[3];
    sub3(va[k], oa, u);
    sub3(vb[m], ob, v);
    for (int q = 0; q < 3; ++q) {
      P[q] = oa[q] + (u[q] * ssc);
      Q[q] = ob[q] + (v[q] * ttc);
    }
  }

  /* vertex-face: j = 0..2 vertex j of B against face A; j = 3..5 vertex j-3
   * of A against face B. */
  double ptmin = 1E30, ss = 0.0, tt = 0.0;
  int id = 0;
  for (int j = 0; j < 6; ++j) {
    const double(*face)[3] = (j < 3) ? va : vb;
    const double* pt = (j < 3) ? vb[j] : va[j - 3];
    double e0[3], e1[3], dd[3];
    sub3(face[1], face[0], e0);
    sub3(face[2], face[0], e1);
    sub3(face[0], pt, dd);
    double s, t;
    double sqr = pt_tri_sqr(OR_DOT(e0, e0), OR_DOT(e0, e1), OR_DOT(e1, e1), OR_DOT(e0, dd),
                            OR_DOT(e1, dd), OR_DOT(dd, dd), &s, &t);
    double dist = sqrt(sqr);
    if (dist < ptmin) {
      ptmin = dist;
      ss = s;
      tt = t;
      id = j;
    }
  }
  if (ptmin > ssmin) return 1;

  const double(*face)[3] = (id < 3) ? va : vb;
  double onface[3];
  for (int q = 0; q < 3; ++q)
    onface[q] = face[0][q] + ((face[2][q] - face[0][q]) * tt) + ((face[1][q] - face[0][q]) * ss);
  if (id < 3) {
    const int vid = (zoltan && id == 1) ? 2 : id;
    for (int q = 0; q < 3; ++q) { P[q] = onface[q]; Q[q] = vb[vid][q]; }
  } else {
    for (int q = 0; q < 3; ++q) { P[q] = va[id - 3][q]; Q[q] = onface[q]; }
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Penalty solver (penalty.cpp:278-431 and 434-581).  Unknowns x = (alpha,
 * beta, gamma, delta) of P = A0 + alpha BA + beta CA, Q = D0 + gamma ED +
 * delta FD; the 4x4 regularised Newton system is solved by the reference's
 * unrolled elimination, in its operation order.
 * ------------------------------------------------------------------------- */
typedef struct pen_ctx {
  double A0[3], B0[3], BA[3], CA[3], ED[3], FD[3];
  double h[16];
  double r;
  double reg0; /* delta before the first step */
} pen_ctx;

static void pen_setup(const double* xA, const double* yA, const double* zA,
                      const double* xB, const double* yB, const double* zB, pen_ctx* c) {
  c->A0[0] = xA[0]; c->A0[1] = yA[0]; c->A0[2] = zA[0];
  c->B0[0] = xB[0]; c->B0[1] = yB[0]; c->B0[2] = zB[0];
  c->BA[0] = xA[1] - xA[0]; c->BA[1] = yA[1] - yA[0]; c->BA[2] = zA[1] - zA[0];
  c->CA[0] = xA[2] - xA[0]; c->CA[1] = yA[2] - yA[0]; c->CA[2] = zA[2] - zA[0];
  c->ED[0] = xB[1] - xB[0]; c->ED[1] = yB[1] - yB[0]; c->ED[2] = zB[1] - zB[0];
  c->FD[0] = xB[2] - xB[0]; c->FD[1] = yB[2] - yB[0]; c->FD[2] = zB[2] - zB[0];
  double* h = c->h;
  h[0] = 2. * OR_DOT(c->BA, c->BA);
  h[1] = 2. * OR_DOT(c->CA, c->BA);
  h[2] = -2. * OR_DOT(c->ED, c->BA);
  h[3] = -2. * OR_DOT(c->FD, c->BA);
  h[4] = h[1];
  h[5] = 2. * OR_DOT(c->CA, c->CA);
  h[6] = -2. * OR_DOT(c->ED, c->CA);
  h[7] = -2. * OR_DOT(c->FD, c->CA);
  h[8] = h[2];
  h[9] = h[6];
  h[10] = 2. * OR_DOT(c->ED, c->ED);
  h[11] = 2. * OR_DOT(c->FD, c->ED);
  h[12] = h[3];
  h[13] = h[7];
  h[14] = h[11];
  h[15] = 2. * OR_DOT(c->FD, c->FD);
  double trace = h[0] + h[5] + h[10] + h[15];
  c->reg0 = trace * 1E-2;
  c->r = sqrt(0.0125 * trace) * 1E5;
}

/* One Newton step: returns dx for the current x and regularisation `reg`. */
static void pen_step(const pen_ctx* c, const double x[4], double reg, double dx[4]) {
  const double* h = c->h;
  const double r = c->r;
  double dh[8], mx[6], sub[3], a[16], b[4];
  int c0 = (-x[0] <= 0), c1 = (-x[1] <= 0), c2 = (x[0] + x[1] - 1 <= 0);
  int c3 = (-x[2] <= 0), c4 = (-x[3] <= 0), c5 = (x[2] + x[3] - 1 <= 0);
  dh[0] = c0 ? 0.0 : -1;
  mx[0] = c0 ? 0.0 : -x[0];
  dh[2] = c1 ? 0.0 : -1;
  mx[1] = c1 ? 0.0 : -x[1];
  dh[1] = dh[3] = c2 ? 0.0 : 1;
  mx[2] = c2 ? 0.0 : x[0] + x[1] - 1;
  dh[4] = c3 ? 0.0 : -1;
  mx[3] = c3 ? 0.0 : -x[2];
  dh[6] = c4 ? 0.0 : -1;
  mx[4] = c4 ? 0.0 : -x[3];
  dh[5] = dh[7] = c5 ? 0.0 : 1;
  mx[5] = c5 ? 0.0 : x[2] + x[3] - 1;
  for (int q = 0; q < 3; ++q)
    sub[q] = (c->A0[q] + (c->BA[q] * x[0]) + (c->CA[q] * x[1])) -
             (c->B0[q] + (c->ED[q] * x[2]) + (c->FD[q] * x[3]));
  double tmp1, tmp2, tmp3, tmp4, tmp5, tmp6;
  b[0] = 2 * OR_DOT(sub, c->BA) + r * (dh[0] * mx[0] + dh[1] * mx[2]);
  a[0] = h[0] + r * (dh[0] * dh[0] + dh[1] * dh[1]) + reg;
  a[4] = h[4] + r * (dh[3] * dh[1]);
  tmp1 = (h[1] + r * (dh[1] * dh[3])) / a[0];
  a[13] = h[13] - h[12] * tmp1;
  a[9] = h[9] - h[8] * tmp1;
  a[5] = (h[5] + r * (dh[2] * dh[2] + dh[3] * dh[3]) + reg) - a[4] * tmp1;
  b[1] = (2 * OR_DOT(sub, c->CA) + r * (dh[2] * mx[1] + dh[3] * mx[2])) - b[0] * tmp1;
  tmp2 = h[2] / a[0];
  tmp3 = h[3] / a[0];
  tmp4 = ((h[6]) - a[4] * tmp2) / a[5];
  a[14] = ((h[14] + r * (dh[7] * dh[5])) - h[12] * tmp2) - a[13] * tmp4;
  a[10] = ((h[10] + r * (dh[4] * dh[4] + dh[5] * dh[5]) + reg) - h[8] * tmp2) - a[9] * tmp4;
  b[2] = ((-2 * OR_DOT(sub, c->ED) + r * (dh[4] * mx[3] + dh[5] * mx[5])) - b[0] * tmp2) -
         b[1] * tmp4;
  tmp5 = (h[7] - a[4] * tmp3) / a[5];
  tmp6 = (((h[11] + r * (dh[5] * dh[7])) - h[8] * tmp3) - a[9] * tmp5) / a[10];
  dx[3] = ((((-2 * OR_DOT(sub, c->FD) + r * (dh[6] * mx[4] + dh[7] * mx[5])) - b[2] * tmp6) -
            b[0] * tmp3) -
           b[1] * tmp5) /
          ((((h[15] + r * (dh[6] * dh[6] + dh[7] * dh[7]) + reg) - h[12] * tmp3) - a[13] * tmp5) -
           a[14] * tmp6);
  dx[2] = (b[2] - (a[14] * dx[3])) / a[10];
  dx[1] = (b[1] - (a[9] * dx[2] + a[13] * dx[3])) / a[5];
  dx[0] = (b[0] - (a[4] * dx[1] + h[8] * dx[2] + h[12] * dx[3])) / a[0];
}

static double dot4(const double a[4], const double b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

static void pen_points(const pen_ctx* c, const double x[4], double P[3], double Q[3]) {
  for (int q = 0; q < 3; ++q) {
    P[q] = c->A0[q] + (c->BA[q] * x[0]) + (c->CA[q] * x[1]);
    Q[q] = c->B0[q] + (c->ED[q] * x[2]) + (c->FD[q] * x[3]);
  }
}

void or_penalty_solver_from(const double* xA, const double* yA, const double* zA,
                            const double* xB, const double* yB, const double* zB,
                            const double x0[4], int iterations,
                            double P[3], double Q[3], double max_err, int* failed,
                            double x_out[4]) {
  pen_ctx c;
  pen_setup(xA, yA, zA, xB, yB, zB, &c);
  double x[4] = {x0[0], x0[1], x0[2], x0[3]};
  double reg = c.reg0;
  int f = 0;
  for (int i = 0; i < iterations; ++i) {
    double dx[4];
    pen_step(&c, x, reg, dx);
    reg = 1E4 * reg;
    f = (dot4(dx, dx) / dot4(x, x) > max_err * max_err) ? 1 : 0;
    for (int q = 0; q < 4; ++q) x[q] -= dx[q];
  }
  if (failed) *failed = f;
  if (x_out) memcpy(x_out, x, sizeof(x));
  pen_points(&c, x, P, Q);
}

void or_penalty_solver(const double* xA, const double* yA, const double* zA,
                       const double* xB, const double* yB, const double* zB,
                       double P[3], double Q[3], double max_err, int* failed) {
  static const double start[4] = {0.33, 0.33, 0.33, 0.33};
  or_penalty_solver_from(xA, yA, zA, xB, yB, zB, start, 4, P, Q, max_err, failed, NULL);
}

void or_penalty_solver_adaptive(const double* xA, const double* yA, const double* zA,
                                const double* xB, const double* yB, const double* zB,
                                double P[3], double Q[3], double max_err, int* its) {
  pen_ctx c;
  pen_setup(xA, yA, zA, xB, yB, zB, &c);
  double x[4] = {0.33, 0.33, 0.33, 0.33};
  double reg = c.reg0;
  for (int i = 0; i < 16; ++i) {
    double dx[4];
    reg = i < 3 ? reg : 1E5 * reg;
    pen_step(&c, x, reg, dx);
    double err = dot4(dx, dx) / dot4(x, x);
    if (err < max_err * max_err) {
      *its = i;
      break;
    }
    for (int q = 0; q < 4; ++q) x[q] = x[q] - dx[q];
  }
  pen_points(&c, x, P, Q);
}

/* ------------------------------------------------------------------ contacts */
static double dist_pq(const double P[3], const double Q[3]) {
  return sqrt(((Q[0] - P[0]) * (Q[0] - P[0])) + ((Q[1] - P[1]) * (Q[1] - P[1])) +
              ((Q[2] - P[2]) * (Q[2] - P[2])));
}

void or_contact_make(double xPA, double yPA, double zPA, double epsA, int idA,
                     double xQB, double yQB, double zQB, double epsB, int idB,
                     int friction, dh_contact* c) {
  memset(c, 0, sizeof(*c));
  c->x[0] = (xPA + xQB) / 2.0;
  c->x[1] = (yPA + yQB) / 2.0;
  c->x[2] = (zPA + zQB) / 2.0;
  c->P[0] = xPA; c->P[1] = yPA; c->P[2] = zPA;
  c->Q[0] = xQB; c->Q[1] = yQB; c->Q[2] = zQB;
  c->eps_total = epsA + epsB;
  double dist = dist_pq(c->P, c->Q);
  c->depth = c->eps_total - dist;
  c->normal[0] = (xPA - xQB) / dist;
  c->normal[1] = (yPA - yQB) / dist;
  c->normal[2] = (zPA - zQB) / dist;
  c->friction = friction ? 1 : 0;
  c->master = idA;
  c->slave = idB;
  c->penetration = 1;
}

/* One row of triangle pairs (triangle iA of A against all of B). */
typedef struct row_buf {
  double* P; /* 3*nB */
  double* Q;
  double* d;
  int* failed;
} row_buf;

static void row_bf(const double* xA, const double* yA, const double* zA, int iA,
                   const double* xB, const double* yB, const double* zB, int nB, row_buf* rb) {
  for (int iB = 0; iB < nB; ++iB) {
    double* P = rb->P + 3 * iB;
    double* Q = rb->Q + 3 * iB;
    or_bf_solver(xA + 3 * iA, yA + 3 * iA, zA + 3 * iA, xB + 3 * iB, yB + 3 * iB, zB + 3 * iB, P, Q);
    rb->d[iB] = dist_pq(P, Q);
  }
}

static void row_penalty(const double* xA, const double* yA, const double* zA, int iA,
                        const double* xB, const double* yB, const double* zB, int nB,
                        double max_err, row_buf* rb) {
  for (int iB = 0; iB < nB; ++iB) {
    double* P = rb->P + 3 * iB;
    double* Q = rb->Q + 3 * iB;
    int f = 0;
    or_penalty_solver(xA + 3 * iA, yA + 3 * iA, zA + 3 * iA, xB + 3 * iB, yB + 3 * iB,
                      zB + 3 * iB, P, Q, max_err, &f);
    rb->failed[iB] = f;
    rb->d[iB] = dist_pq(P, Q);
  }
}

typedef struct emit_ctx {
  dh_contact* out;
  int cap;
  int n;
  double epsA, epsB;
  int idA, idB, fric;
} emit_ctx;

static void emit(emit_ctx* e, const double P[3], const double Q[3], int iA, int iB) {
  if (e->n < e->cap && e->out) {
    dh_contact* c = e->out + e->n;
    or_contact_make(P[0], P[1], P[2], e->epsA, e->idA, Q[0], Q[1], Q[2], e->epsB, e->idB,
                    e->fric, c);
    c->triA = iA;
    c->triB = iB;
    c->pair = 0;
  }
  e->n++;
}

int or_detect(int method,
              const double* xA, const double* yA, const double* zA, int nA,
              double epsA, int fricA, int idA,
              const double* xB, const double* yB, const double* zB, int nB,
              double epsB, int fricB, int idB,
              dh_contact* out, int cap, int* n_out, dh_stats* stats) {
  if (nA < 0 || nB < 0 || !n_out) return DH_ERR_ARG;
  if (method == DH_HYBRID_TRIANGLE && nB > nA) return DH_ERR_REF_UB;
  dh_stats dummy;
  if (!stats) {
    memset(&dummy, 0, sizeof(dummy));
    stats = &dummy;
  }
  emit_ctx e = {out, cap, 0, epsA, epsB, idA, idB, (fricA && fricB) ? 1 : 0};
  const double margin = epsA + epsB;
  row_buf rb;
  rb.P = (double*)malloc(sizeof(double) * 3 * (nB + 1));
  rb.Q = (double*)malloc(sizeof(double) * 3 * (nB + 1));
  rb.d = (double*)malloc(sizeof(double) * (nB + 1));
  rb.failed = (int*)calloc(nB + 1, sizeof(int));
  stats->triangle_checks += (int64_t)nA * nB;

  switch (method) {
    case DH_BF:
      for (int iA = 0; iA < nA; ++iA) {
        row_bf(xA, yA, zA, iA, xB, yB, zB, nB, &rb);
        stats->pairs_evaluated += nB;
        for (int iB = 0; iB < nB; ++iB)
          if (rb.d[iB] <= margin) emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
      }
      break;
    case DH_PENALTY: {
      const double max_err = (epsA + epsB) / 16.0;
      for (int iA = 0; iA < nA; ++iA) {
        row_penalty(xA, yA, zA, iA, xB, yB, zB, nB, max_err, &rb);
        stats->pairs_evaluated += nB;
        for (int iB = 0; iB < nB; ++iB)
          if (rb.d[iB] < margin) emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
      }
    } break;
    case DH_PENALTY_STAT: {
      /* penalty.cpp:50-111: adaptive solver, iteration histogram, and an
       * emission loop that only visits iB < nB in point units (no *3). */
      for (int iA = 0; iA < nA; ++iA) {
        const double max_err = (epsA + epsB) / 16.0;
        for (int iB = 0; iB < nB; ++iB) {
          int its = 0;
          or_penalty_solver_adaptive(xA + 3 * iA, yA + 3 * iA, zA + 3 * iA, xB + 3 * iB,
                                     yB + 3 * iB, zB + 3 * iB, rb.P + 3 * iB, rb.Q + 3 * iB,
                                     max_err, &its);
          stats->newton_histogram[its]++;
          rb.d[iB] = dist_pq(rb.P + 3 * iB, rb.Q + 3 * iB);
        }
        stats->pairs_evaluated += nB;
        for (int iBp = 0; iBp < nB; iBp += 3) {
          int iB = iBp / 3;
          if (rb.d[iB] <= margin) emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
        }
      }
    } break;
    case DH_HYBRID_TRIANGLE: {
      /* hybrid.cpp:59-200: rows run to nB (the reference's loop bound). */
      const double max_err = (epsA + epsB) / 16;
      for (int iA = 0; iA < nB; ++iA) {
        row_penalty(xA, yA, zA, iA, xB, yB, zB, nB, max_err, &rb);
        stats->pairs_evaluated += nB;
        int any = 0;
        for (int iB = 0; iB < nB; ++iB) any |= rb.failed[iB];
        if (any) {
          row_bf(xA, yA, zA, iA, xB, yB, zB, nB, &rb);
          stats->pairs_evaluated += nB;
          stats->fallback_rows++;
        }
        for (int iB = 0; iB < nB; ++iB)
          if (rb.d[iB] < margin) emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
      }
    } break;
    case DH_HYBRID_BATCH: {
      const double max_err = (epsA + epsB) / 16;
      for (int iA = 0; iA < nA; ++iA) {
        row_penalty(xA, yA, zA, iA, xB, yB, zB, nB, max_err, &rb);
        stats->pairs_evaluated += nB;
        int counter = 0;
        for (int iB = 0; iB < nB; ++iB)
          if (rb.failed[iB]) counter++;
        double local_err = counter * max_err;
        if (local_err > ((double)max_err * (double)(3 * nB)) / 32.0) {
          row_bf(xA, yA, zA, iA, xB, yB, zB, nB, &rb);
          stats->pairs_evaluated += nB;
          stats->fallback_rows++;
        }
        for (int iB = 0; iB < nB; ++iB)
          if (rb.d[iB] < margin) emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
      }
    } break;
    case DH_HYBRID_BATCH_STAT:
    case DH_HYBRID_TRIANGLE_STAT: {
      /* hybrid.cpp:202-336 and 489-627 (identical arithmetic): fallback when
       * count*maxErr > maxErr*nB/32; only the row minimum is emitted. */
      for (int iA = 0; iA < nA; ++iA) {
        const double max_err = (epsA + epsB) / 16;
        stats->batch_size += nB;
        stats->batch_error += max_err;
        row_penalty(xA, yA, zA, iA, xB, yB, zB, nB, max_err, &rb);
        stats->pairs_evaluated += nB;
        int counter = 0;
        for (int iB = 0; iB < nB; ++iB)
          if (rb.failed[iB]) counter++;
        double local_err = counter * max_err;
        stats->penalty_fails += counter;
        if (local_err > ((double)max_err * (double)nB) / 32) {
          stats->batch_fails += 1;
          row_bf(xA, yA, zA, iA, xB, yB, zB, nB, &rb);
          stats->pairs_evaluated += nB;
          stats->fallback_rows++;
        }
        double min_d = 1E99;
        for (int iB = 0; iB < nB; ++iB) min_d = (min_d < rb.d[iB]) ? min_d : rb.d[iB];
        for (int iB = 0; iB < nB; ++iB)
          if (rb.d[iB] <= min_d && rb.d[iB] < margin)
            emit(&e, rb.P + 3 * iB, rb.Q + 3 * iB, iA, iB);
      }
    } break;
    case DH_GJK: {
      /* gjk.cpp:25-100 over every vertex of both particles (the reference
       * copies nTri of them and then reads 3*nTri points from that buffer,
       * past its end); one contact when |Q-P| <= epsA + epsB */
      double* a = (double*)malloc(sizeof(double) * 9 * (nA > 0 ? nA : 1));
      double* b = (double*)malloc(sizeof(double) * 9 * (nB > 0 ? nB : 1));
      double P[3], Q[3];
      for (int i = 0; i < 3 * nA; ++i) { a[3 * i] = xA[i]; a[3 * i + 1] = yA[i]; a[3 * i + 2] = zA[i]; }
      for (int i = 0; i < 3 * nB; ++i) { b[3 * i] = xB[i]; b[3 * i + 1] = yB[i]; b[3 * i + 2] = zB[i]; }
      if (nA > 0 && nB > 0) {
        const double g = or_gjk(a, 3 * nA, b, 3 * nB, P, Q);
        stats->pairs_evaluated += 1;
        const double dist = sqrt(((Q[0] - P[0]) * (Q[0] - P[0])) + ((Q[1] - P[1]) * (Q[1] - P[1])) +
                                 ((Q[2] - P[2]) * (Q[2] - P[2])));
        if (g == g && dist <= margin) emit(&e, P, Q, -1, -1);
      }
      free(a); free(b);
    } break;
    default:
      free(rb.P); free(rb.Q); free(rb.d); free(rb.failed);
      return DH_ERR_ARG;
  }
  free(rb.P); free(rb.Q); free(rb.d); free(rb.failed);
  *n_out = e.n;
  return (e.n > cap) ? DH_ERR_CAPACITY : DH_OK;
}

static int zh_segt(const double p1[3], const double p2[3], const double A[3], const double B[3],
                   const double C[3], double P[3]);

/* The intersection pre-test of the exact mode: the six segment-triangle tests
 * of zoltan/hybrid.cpp:652-670 in their order; P = the crossing point. */
static int segt_chain(const double va[3][3], const double vb[3][3], double P[3]) {
  if (zh_segt(vb[0], vb[1], va[0], va[1], va[2], P) == 1) return 1;
  if (zh_segt(vb[1], vb[2], va[0], va[1], va[2], P) == 1) return 1;
  if (zh_segt(vb[2], vb[0], va[0], va[1], va[2], P) == 1) return 1;
  if (zh_segt(va[0], va[1], vb[0], vb[1], vb[2], P) == 1) return 1;
  if (zh_segt(va[1], va[2], vb[0], vb[1], vb[2], P) == 1) return 1;
  if (zh_segt(va[2], va[0], vb[0], vb[1], vb[2], P) == 1) return 1;
  return 0;
}

int or_tri_pairs(int method, const double* soa, int64_t n, double eps_total,
                 double* P, double* Q, double* d, uint8_t* flags) {
  const double max_err = eps_total / 16.0;
  for (int64_t i = 0; i < n; ++i) {
    double xa[3], ya[3], za[3], xb[3], yb[3], zb[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
      xa[k] = soa[(0 + k) * n + i];
      ya[k] = soa[(3 + k) * n + i];
      za[k] = soa[(6 + k) * n + i];
      xb[k] = soa[(9 + k) * n + i];
      yb[k] = soa[(12 + k) * n + i];
      zb[k] = soa[(15 + k) * n + i];
    }
    uint8_t fl = 0;
    switch (method) {
      case DH_PAIR_BF64:
        if (or_bf_solver(xa, ya, za, xb, yb, zb, p, q)) fl |= DH_FLAG_EDGE;
        break;
      case DH_PAIR_PENALTY64: {
        int f = 0;
        or_penalty_solver(xa, ya, za, xb, yb, zb, p, q, max_err, &f);
        if (f) fl |= DH_FLAG_FAILED;
      } break;
      case DH_PAIR_HYBRID64: {
        int f = 0;
        or_penalty_solver(xa, ya, za, xb, yb, zb, p, q, max_err, &f);
        if (f) {
          fl |= DH_FLAG_FAILED | DH_FLAG_RECHECK;
          if (or_bf_solver(xa, ya, za, xb, yb, zb, p, q)) fl |= DH_FLAG_EDGE;
        }
      } break;
      case DH_PAIR_PENALTY_ADAPT64: {
        int its = 0;
        or_penalty_solver_adaptive(xa, ya, za, xb, yb, zb, p, q, max_err, &its);
      } break;
      case DH_PAIR_BF64_EXACT: {
        double va[3][3], vb[3][3], x[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
          va[k][0] = xa[k]; va[k][1] = ya[k]; va[k][2] = za[k];
          vb[k][0] = xb[k]; vb[k][1] = yb[k]; vb[k][2] = zb[k];
        }
        if (segt_chain(va, vb, x)) {
          for (int c = 0; c < 3; ++c) p[c] = q[c] = x[c];
          fl |= DH_FLAG_INTERSECT;
        } else if (or_bf_solver(xa, ya, za, xb, yb, zb, p, q)) {
          fl |= DH_FLAG_EDGE;
        }
      } break;
      default:
        return DH_ERR_ARG;
    }
    for (int c = 0; c < 3; ++c) {
      P[c * n + i] = p[c];
      Q[c * n + i] = q[c];
    }
    d[i] = dist_pq(p, q);
    if (flags) flags[i] = fl;
  }
  return DH_OK;
}

/* -------------------------------------------------------------------- forces */
void or_contacts_forces(const dh_contact* cs, int nc, int negate_normal,
                        const double posA[3], const double angA[3], const double linA[3],
                        double massA,
                        const double posB[3], const double angB[3], const double linB[3],
                        double massB,
                        const dh_force_params* fp, double force[3], double torque[3]) {
  for (int k = 0; k < nc; ++k) {
    const dh_contact* c = cs + k;
    double n[3];
    for (int q = 0; q < 3; ++q) n[q] = negate_normal ? -1.0 * c->normal[q] : c->normal[q];
    double z[3], vi[3], vj[3], vij[3];
    sub3(c->x, posA, z);
    vi[0] = angA[1] * z[2] - angA[2] * z[1] + linA[0];
    vi[1] = angA[2] * z[0] - angA[0] * z[2] + linA[1];
    vi[2] = angA[0] * z[1] - angA[1] * z[0] + linA[2];
    sub3(c->x, posB, z);
    vj[0] = angB[1] * z[2] - angB[2] * z[1] + linB[0];
    vj[1] = angB[2] * z[0] - angB[0] * z[2] + linB[1];
    vj[2] = angB[0] * z[1] - angB[1] * z[0] + linB[2];
    sub3(vj, vi, vij);

    /* spring (forces.cpp:208-220) / springSphere (227-249) */
    double ma = 1.0 / ((1.0 / massA) + (1.0 / massB));
    double vel = (vij[0] * n[0]) + (vij[1] * n[1]) + (vij[2] * n[2]);
    double damp = fp->damper * 2.0 * sqrt(ma * fp->spring) * vel;
    double fn = fp->spring * c->depth + damp;
    double f[3] = {fn * n[0], fn * n[1], fn * n[2]};
    double fr[3] = {0.0, 0.0, 0.0};
    if (c->friction) {
      /* friction (forces.cpp:255-289): tangential part of v_i, not v_ij. */
      double vt[3];
      for (int q = 0; q < 3; ++q)
        vt[q] = vi[q] - n[q] * ((vi[0] * n[0]) + (vi[1] * n[1]) + (vi[2] * n[2]));
      for (int q = 0; q < 3; ++q) fr[q] = -vt[q] * fp->friction * fn;
    }
    for (int q = 0; q < 3; ++q) force[q] += f[q] + fr[q];
    double arm[3];
    sub3(c->x, posA, arm);
    torque[0] += arm[1] * (f[2]) - arm[2] * (f[1]);
    torque[1] += arm[2] * (f[0]) - arm[0] * (f[2]);
    torque[2] += arm[0] * (f[1]) - arm[1] * (f[0]);
  }
}

/* ------------------------------------------------------------------ dynamics */
void or_expmap(double o1, double o2, double o3, double L[9]) {
  double v0 = o1 * o1, v1 = o2 * o2, v2 = o3 * o3;
  double angsq = v0 + v1 + v2;
  double sx, cx;
  if (angsq < 3.0461741978671E-02) {
    /* Taylor series of sin(x)/x and (1-cos x)/x^2 in angsq = x^2 (1/(2k+1)!
     * and 1/(2k+2)!, rounded to 16 digits as in dynamics.cpp:41-64). */
    sx = 1.0 + (-1.666666666666667E-1 +
                (8.333333333333333E-3 +
                 (-1.984126984126984E-4 +
                  (2.755731922398589E-6 +
                   (-2.505210838544172E-8 + 1.605904383682161E-10 * angsq) * angsq) *
                      angsq) *
                     angsq) *
                    angsq) *
                   angsq;
    cx = 0.5 + (-4.166666666666667E-2 +
                (1.388888888888889E-3 +
                 (-2.480158730158730E-5 +
                  (2.755731922398589E-7 +
                   (-2.087675698786810E-9 + 1.147074559772972E-11 * angsq) * angsq) *
                      angsq) *
                     angsq) *
                    angsq) *
                   angsq;
  } else {
    double t = angsq;
    double ang = sqrt(angsq);
    sx = sin(ang) / ang;
    cx = (1.0 - cos(ang)) / t;
  }
  double v01 = o1 * o2;
  double v02 = o2 * o3; /* the reference's product (dynamics.cpp:78) */
  double v12 = o2 * o3;
  double s0 = sx * o1, s1 = sx * o2, s2 = sx * o3;
  L[0] = -cx * (v2 + v1);
  L[1] = cx * v01;
  L[2] = cx * v02;
  L[3] = L[1];
  L[4] = -cx * (v2 + v0);
  L[5] = cx * v12;
  L[6] = L[2];
  L[7] = L[5];
  L[8] = -cx * (v1 + v0);
  L[0] += 1.0;
  L[1] += s2;
  L[2] -= s1;
  L[3] -= s2;
  L[4] += 1.0;
  L[5] += s0;
  L[6] += s1;
  L[7] -= s0;
  L[8] += 1.0;
}

void or_update_rotation_matrix(double angular[3], const double refAngular[3],
                               double rotation[9], double step) {
  double DL[9], r0[9];
  or_expmap(step * refAngular[0], step * refAngular[1], step * refAngular[2], DL);
  memcpy(r0, rotation, sizeof(r0));
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row)
      rotation[3 * col + row] =
          r0[row] * DL[3 * col] + r0[3 + row] * DL[3 * col + 1] + r0[6 + row] * DL[3 * col + 2];
  for (int row = 0; row < 3; ++row)
    angular[row] = rotation[row] * refAngular[0] + rotation[3 + row] * refAngular[1] +
                   rotation[6 + row] * refAngular[2];
}

void or_update_angular(double refAngular[3], const double rotation[9],
                       const double inertia[9], const double inverse[9],
                       const double torque[3], double step) {
  double half = 0.5 * step;
  double T[3], DL[9], A[3], B[3];
  for (int row = 0; row < 3; ++row)
    T[row] = rotation[3 * row] * torque[0] + rotation[3 * row + 1] * torque[1] +
             rotation[3 * row + 2] * torque[2];
  or_expmap(-half * refAngular[0], -half * refAngular[1], -half * refAngular[2], DL);
  for (int r = 0; r < 3; ++r)
    A[r] = inertia[r] * refAngular[0] + inertia[3 + r] * refAngular[1] + inertia[6 + r] * refAngular[2];
  for (int r = 0; r < 3; ++r) B[r] = DL[r] * A[0] + DL[3 + r] * A[1] + DL[6 + r] * A[2];
  for (int r = 0; r < 3; ++r) B[r] = B[r] + half * T[r];
  for (int r = 0; r < 3; ++r) A[r] = inverse[r] * B[0] + inverse[3 + r] * B[1] + inverse[6 + r] * B[2];
  for (int r = 0; r < 3; ++r) B[r] = inertia[r] * A[0] + inertia[3 + r] * A[1] + inertia[6 + r] * A[2];
  T[0] -= A[1] * B[2] - A[2] * B[1];
  T[1] -= A[2] * B[0] - A[0] * B[2];
  T[2] -= A[0] * B[1] - A[1] * B[0];
  for (int r = 0; r < 3; ++r) T[r] *= step;
  for (int r = 0; r < 3; ++r)
    refAngular[r] = refAngular[r] + inverse[r] * T[0] + inverse[3 + r] * T[1] + inverse[6 + r] * T[2];
}

void or_update_vertex(double* x, double* y, double* z, double rx, double ry, double rz,
                      const double R[9], const double pos[3], const double refpos[3]) {
  double C[3] = {rx - refpos[0], ry - refpos[1], rz - refpos[2]};
  *x = pos[0] + (R[0] * C[0] + R[3] * C[1] + R[6] * C[2]);
  *y = pos[1] + (R[1] * C[0] + R[4] * C[1] + R[7] * C[2]);
  *z = pos[2] + (R[2] * C[0] + R[5] * C[1] + R[8] * C[2]);
}

/* -------------------------------------------------------------------- step */
static void particle_aabb(const or_scene* s, int p, double lo[3], double hi[3]) {
  lo[0] = lo[1] = lo[2] = INFINITY;
  hi[0] = hi[1] = hi[2] = -INFINITY;
  for (int v = 3 * s->tri_offset[p]; v < 3 * s->tri_offset[p + 1]; ++v) {
    double c[3] = {s->x[v], s->y[v], s->z[v]};
    for (int q = 0; q < 3; ++q) {
      if (c[q] < lo[q]) lo[q] = c[q];
      if (c[q] > hi[q]) hi[q] = c[q];
    }
  }
  /* inflate by the margin plus a 1e-9 m band, so that a pair whose true
   * distance is within eps_total is never culled */
  for (int q = 0; q < 3; ++q) {
    lo[q] -= s->eps[p] + 0.5e-9;
    hi[q] += s->eps[p] + 0.5e-9;
  }
}

int64_t or_step(or_scene* s, int method, const dh_force_params* fp, double dt,
                const double gravity[3], int64_t* checks, int* npairs_out) {
  const int n = s->n;
  double* lo = (double*)malloc(sizeof(double) * 3 * n);
  double* hi = (double*)malloc(sizeof(double) * 3 * n);
  for (int p = 0; p < n; ++p) particle_aabb(s, p, lo + 3 * p, hi + 3 * p);
  double* force = (double*)calloc(3 * n, sizeof(double));
  double* torque = (double*)calloc(3 * n, sizeof(double));
  int64_t total_contacts = 0;
  int npairs = 0;
  int cap = 4096;
  dh_contact* buf = (dh_contact*)malloc(sizeof(dh_contact) * cap);
  /* Per-pair force sums are stored first and then added to each particle in
   * ascending partner index. */
  typedef struct pr { int i, j; double fA[3], tA[3], fB[3], tB[3]; } pr;
  int prcap = 1024, npr = 0;
  pr* prs = (pr*)malloc(sizeof(pr) * prcap);
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      if (s->b.obstacle[i] && s->b.obstacle[j]) continue;
      int ov = 1;
      for (int q = 0; q < 3; ++q)
        if (lo[3 * i + q] > hi[3 * j + q] || lo[3 * j + q] > hi[3 * i + q]) ov = 0;
      if (!ov) continue;
      npairs++;
      int oi = s->tri_offset[i], oj = s->tri_offset[j];
      int ni = s->tri_offset[i + 1] - oi, nj = s->tri_offset[j + 1] - oj;
      if (checks) *checks += (int64_t)ni * nj;
      int nc = 0;
      int rc;
      for (;;) {
        rc = or_detect(method, s->x + 3 * oi, s->y + 3 * oi, s->z + 3 * oi, ni, s->eps[i],
                       s->friction[i], s->id[i], s->x + 3 * oj, s->y + 3 * oj, s->z + 3 * oj,
                       nj, s->eps[j], s->friction[j], s->id[j], buf, cap, &nc, NULL);
        if (rc != DH_ERR_CAPACITY) break;
        cap = nc * 2;
        buf = (dh_contact*)realloc(buf, sizeof(dh_contact) * cap);
      }
      if (nc == 0) continue;
      total_contacts += nc;
      if (npr == prcap) {
        prcap *= 2;
        prs = (pr*)realloc(prs, sizeof(pr) * prcap);
      }
      pr* r = prs + npr++;
      memset(r, 0, sizeof(*r));
      r->i = i;
      r->j = j;
      const dh_bodies* b = &s->b;
      or_contacts_forces(buf, nc, 0, b->pos + 3 * i, b->angvel + 3 * i, b->linvel + 3 * i,
                         b->mass[i], b->pos + 3 * j, b->angvel + 3 * j, b->linvel + 3 * j,
                         b->mass[j], fp, r->fA, r->tA);
      or_contacts_forces(buf, nc, 1, b->pos + 3 * j, b->angvel + 3 * j, b->linvel + 3 * j,
                         b->mass[j], b->pos + 3 * i, b->angvel + 3 * i, b->linvel + 3 * i,
                         b->mass[i], fp, r->fB, r->tB);
    }
  }
  /* particle p: partners ascending.  Pairs with p as the second member
   * (partner i < p) come before pairs with p first (partner j > p).
   * A particle without any contact is skipped by the force/angular update
   * (Engine.cpp:464, Collision.cpp:1056): updateAngular is not a no-op for
   * zero torque, so the skip is observable. */
  int* touched = (int*)calloc(n, sizeof(int));
  for (int k = 0; k < npr; ++k) touched[prs[k].i] = touched[prs[k].j] = 1;
  for (int p = 0; p < n; ++p) {
    for (int k = 0; k < npr; ++k)
      if (prs[k].j == p)
        for (int q = 0; q < 3; ++q) {
          force[3 * p + q] += prs[k].fB[q];
          torque[3 * p + q] += prs[k].tB[q];
        }
    for (int k = 0; k < npr; ++k)
      if (prs[k].i == p)
        for (int q = 0; q < 3; ++q) {
          force[3 * p + q] += prs[k].fA[q];
          torque[3 * p + q] += prs[k].tA[q];
        }
  }
  dh_bodies* b = &s->b;
  for (int p = 0; p < n; ++p) {
    if (b->obstacle[p] || !touched[p]) continue;
    for (int q = 0; q < 3; ++q) b->linvel[3 * p + q] += dt * (force[3 * p + q] / b->mass[p]);
    or_update_angular(b->refangvel + 3 * p, b->rot + 9 * p, b->inertia + 9 * p,
                      b->inverse + 9 * p, torque + 3 * p, dt);
  }
  for (int p = 0; p < n; ++p) {
    /* obstacles at rest are skipped (Engine.cpp:532); one with a velocity
     * moves with it kinematically (no force, no gravity) */
    const int kinematic = b->obstacle[p] != 0;
    if (kinematic) {
      const double* u = b->linvel + 3 * p;
      const double* W = b->refangvel + 3 * p;
      if (u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0 && W[0] == 0.0 && W[1] == 0.0 && W[2] == 0.0)
        continue;
    }
    for (int q = 0; q < 3; ++q)
      if (gravity[q] != 0.0 && !kinematic) b->linvel[3 * p + q] += dt * gravity[q];
    for (int q = 0; q < 3; ++q) b->pos[3 * p + q] += dt * b->linvel[3 * p + q];
    or_update_rotation_matrix(b->angvel + 3 * p, b->refangvel + 3 * p, b->rot + 9 * p, dt);
    for (int v = 3 * s->tri_offset[p]; v < 3 * s->tri_offset[p + 1]; ++v)
      or_update_vertex(s->x + v, s->y + v, s->z + v, s->rx[v], s->ry[v], s->rz[v],
                       b->rot + 9 * p, b->pos + 3 * p, b->refpos + 3 * p);
  }
  free(lo); free(hi); free(force); free(torque); free(buf); free(prs); free(touched);
  if (npairs_out) *npairs_out = npairs;
  return total_contacts;
}

/* ------------------------------------------------------------------ GJK */
/* gjk.cpp:102-460: Gilbert-Johnson-Keerthi distance of two convex point sets
 * (AoS, 3 doubles per point), Johnson's recursive sub-simplex projection. */
#define GJK_EPS 1e-6 /* GEOMETRIC_EPSILON, gjk.cpp:102 */

static double g_dot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* sub-simplex bit sets in trial order and the trial counts (gjk.cpp:126-148) */
static const int g_sets[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 11, 13, 14, 15};
static const int g_tries[5] = {0, 0, 3, 7, 15};

typedef struct { double w[3]; const double* a; const double* b; } g_vert;

/* gjk.cpp:302-414 */
static int g_project(g_vert* s, int n, double* l, double* v) {
  double dt[4][4], cof[16][4];
  if (n == 1) {
    l[0] = 1.0;
    v[0] = s[0].w[0]; v[1] = s[0].w[1]; v[2] = s[0].w[2];
    return 1;
  }
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) dt[i][j] = dt[j][i] = g_dot(s[i].w, s[j].w);
  for (int m = 0; m < g_tries[n]; ++m) {
    const int set = g_sets[m];
    int good = 1, first = -1;
    int members = 0;
    for (int i = 0; i < n; ++i) members += (set >> i) & 1;
    if (members == 1) {
      for (int i = 0; i < n; ++i)
        if (set & (1 << i)) { first = i; break; }
      cof[set][first] = 1.0;
      for (int j = 0; j < n; ++j) {
        if (set & (1 << j)) continue;
        cof[set][j] = dt[first][first] - dt[first][j];
        if (cof[set][j] > 0.0) good = 0;
      }
    } else {
      for (int i = 0; i < n; ++i) {
        if (!(set & (1 << i))) continue;
        if (first < 0) first = i;
        cof[set][i] = cof[set & ~(1 << i)][i];
        if (cof[set][i] <= 0.0) good = 0;
      }
      for (int j = 0; j < n; ++j) {
        if (set & (1 << j)) continue;
        cof[set][j] = 0.0;
        for (int i = 0; i < n; ++i)
          if (set & (1 << i)) cof[set][j] += cof[set][i] * (dt[i][first] - dt[i][j]);
        if (cof[set][j] > 0.0) good = 0;
      }
    }
    if (good) {
      double sum = 0.0;
      int o = 0;
      for (int i = 0; i < n; ++i) {
        if (!(set & (1 << i))) continue;
        s[o] = s[i];
        l[o] = cof[set][i];
        sum += l[o];
        ++o;
      }
      v[0] = v[1] = v[2] = 0.0;
      for (int i = 0; i < o; ++i) {
        l[i] /= sum;
        v[0] = v[0] + l[i] * s[i].w[0];
        v[1] = v[1] + l[i] * s[i].w[1];
        v[2] = v[2] + l[i] * s[i].w[2];
      }
      return o;
    }
  }
  return 0;
}

static const double* g_support(const double* c, int n, const double* v, int maximal) {
  const double* best = NULL;
  double bd = maximal ? -DBL_MAX : DBL_MAX;
  for (int i = 0; i < n; ++i) {
    const double d = g_dot(c + 3 * i, v);
    if (maximal ? (d > bd) : (d < bd)) { bd = d; best = c + 3 * i; }
  }
  return best;
}

double or_gjk(const double* a, int na, const double* b, int nb, double p[3], double q[3]) {
  g_vert s[4];
  double v[3], l[4], mi = 0.0, vlen;
  int far = 1, n = 0, it = 0;
  const long long lim = (long long)(na + nb) * (na + nb);
  v[0] = a[0] - b[0]; v[1] = a[1] - b[1]; v[2] = a[2] - b[2];
  vlen = sqrt(g_dot(v, v));
  while (far && vlen > GJK_EPS && n < 4 && it++ < lim) {
    s[n].a = g_support(a, na, v, 0);
    s[n].b = g_support(b, nb, v, 1);
    if (!s[n].a || !s[n].b) return NAN;
    for (int c = 0; c < 3; ++c) s[n].w[c] = s[n].a[c] - s[n].b[c];
    {
      const double delta = g_dot(v, s[n].w) / vlen;
      mi = mi > delta ? mi : delta;
    }
    far = (vlen - mi) > GJK_EPS;
    if (far) {
      n = g_project(s, n + 1, l, v);
      vlen = sqrt(g_dot(v, v));
    }
  }
  if (n) {
    p[0] = p[1] = p[2] = 0.0;
    q[0] = q[1] = q[2] = 0.0;
    for (--n; n >= 0; --n)
      for (int c = 0; c < 3; ++c) {
        p[c] = p[c] + l[n] * s[n].a[c];
        q[c] = q[c] + l[n] * s[n].b[c];
      }
  } else {
    for (int c = 0; c < 3; ++c) { p[c] = a[c]; q[c] = b[c]; }
  }
  return vlen;
}

/* ---------------------------------------------------------------------------
 * zoltan application kernels (zoltan/penalty.cpp, zoltan/hybrid.cpp).
 * ------------------------------------------------------------------------- */
static void tri_tables(const double a[3], const double b[3], const double c[3],
                       double x[3], double y[3], double z[3]) {
  x[0] = a[0]; y[0] = a[1]; z[0] = a[2];
  x[1] = b[0]; y[1] = b[1]; z[1] = b[2];
  x[2] = c[0]; y[2] = c[1]; z[2] = c[2];
}

/* zoltan/penalty.cpp:27-208 */
void or_zoltan_penalty(unsigned it, unsigned nt, const double a[3], const double b[3],
                       const double c[3], double* const d[3], double* const e[3],
                       double* const f[3], double* const p[3], double* const q[3]) {
  double xa[3], ya[3], za[3];
  tri_tables(a, b, c, xa, ya, za);
  for (unsigned i = it; i < nt; ++i) {
    double xb[3] = {d[0][i], e[0][i], f[0][i]}, yb[3] = {d[1][i], e[1][i], f[1][i]},
           zb[3] = {d[2][i], e[2][i], f[2][i]};
    pen_ctx cx;
    pen_setup(xa, ya, za, xb, yb, zb, &cx);
    const double trace = cx.h[0] + cx.h[5] + cx.h[10] + cx.h[15];
    double delta = trace * 0.1;
    const double lambda = sqrt(0.0125 * trace);
    cx.r = lambda * pow(10, log(lambda) + 14);
    double x[4] = {0.33, 0.33, 0.33, 0.33};
    for (int j = 0; j < 5; ++j) {
      double dx[4];
      delta = i < 4 ? delta : cx.r * delta;
      pen_step(&cx, x, delta, dx);
      for (int k = 0; k < 4; ++k) x[k] = x[k] - dx[k];
    }
    double P[3], Q[3];
    pen_points(&cx, x, P, Q);
    for (int k = 0; k < 3; ++k) { p[k][i] = P[k]; q[k][i] = Q[k]; }
  }
}

/* zoltan/hybrid.cpp:134-278: 7 steps, error of the last one */
static void zh_penalty(const double A[3], const double B[3], const double C[3],
                       const double D[3], const double E[3], const double F[3],
                       double P[3], double Q[3], double* error) {
  double xa[3], ya[3], za[3], xb[3], yb[3], zb[3];
  tri_tables(A, B, C, xa, ya, za);
  tri_tables(D, E, F, xb, yb, zb);
  pen_ctx cx;
  pen_setup(xa, ya, za, xb, yb, zb, &cx);
  const double trace = cx.h[0] + cx.h[5] + cx.h[10] + cx.h[15];
  const double delta = trace * 1E-16;
  const double lambda = sqrt(0.0125 * trace);
  cx.r = lambda * pow(10, log10(lambda) + 8);
  double x[4] = {0.33, 0.33, 0.33, 0.33}, dx[4] = {0, 0, 0, 0};
  for (int i = 0; i < 7; ++i) {
    pen_step(&cx, x, delta, dx);
    for (int k = 0; k < 4; ++k) x[k] = x[k] - dx[k];
  }
  *error = sqrt(dot4(dx, dx) / dot4(x, x));
  pen_points(&cx, x, P, Q);
}

/* zoltan/hybrid.cpp:280-342 */
static int zh_segt(const double p1[3], const double p2[3], const double A[3], const double B[3],
                   const double C[3], double P[3]) {
  double u[3], v[3], n[3], dir[3], w0[3], w[3];
  sub3(B, A, u);
  sub3(C, A, v);
  n[0] = u[1] * v[2] - u[2] * v[1];
  n[1] = u[2] * v[0] - u[0] * v[2];
  n[2] = u[0] * v[1] - u[1] * v[0];
  if (n[0] == 0 && n[1] == 0 && n[2] == 0) return -1;
  sub3(p2, p1, dir);
  sub3(p1, A, w0);
  const double a = -OR_DOT(n, w0);
  const double b = OR_DOT(n, dir);
  if (fabs(b) < 1E-99 && a == 0) return 2;
  const double r = a / b;
  if (r < 0.0 || r > 1.0) return 0;
  for (int k = 0; k < 3; ++k) P[k] = p1[k] + r * dir[k];
  const double uu = OR_DOT(u, u), uv = OR_DOT(u, v), vv = OR_DOT(v, v);
  sub3(P, A, w);
  const double wu = OR_DOT(w, u), wv = OR_DOT(w, v);
  const double D = uv * uv - uu * vv;
  const double s = (uv * wv - vv * wu) / D;
  if (s < 0.0 || s > 1.0) return 0;
  const double t = (uv * wu - uu * wv) / D;
  if (t < 0.0 || (s + t) > 1.0) return 0;
  return 1;
}

/* zoltan/hybrid.cpp:344-445, including Q's y and z taken from p4's x */
static double zh_segseg(const double p1[3], const double p2[3], const double p3[3],
                        const double p4[3], double P[3], double Q[3]) {
  double u[3], v[3], w[3];
  sub3(p1, p2, u);
  sub3(p3, p4, v);
  sub3(p2, p4, w);
  const double a = OR_DOT(u, u), b = OR_DOT(u, v), c = OR_DOT(v, v), d = OR_DOT(u, w),
               e = OR_DOT(v, w);
  const double D = a * c - b * b;
  double sD = D, tD = D, sN = 0, tN = 0;
  const double small = 0.00000001;
  if (D < small) {
    sN = 0.0; sD = 1.0; tN = e; tD = c;
  } else {
    sN = (b * e - c * d);
    tN = (a * e - b * d);
    if (sN < 0.0) { sN = 0.0; tN = e; tD = c; }
    else if (sN > sD) { sN = sD; tN = e + b; tD = c; }
  }
  if (tN < 0.0) {
    tN = 0.0;
    if (-d < 0.0) sN = 0.0;
    else if (-d > a) sN = sD;
    else { sN = -d; sD = a; }
  } else if (tN > tD) {
    tN = tD;
    if ((-d + b) < 0.0) sN = 0;
    else if ((-d + b) > a) sN = sD;
    else { sN = (-d + b); sD = a; }
  }
  const double sc = (fabs(sN) < small) ? 0.0 : sN / sD;
  const double tc = (fabs(tN) < small) ? 0.0 : tN / tD;
  for (int k = 0; k < 3; ++k) P[k] = p2[k] + u[k] * sc;
  Q[0] = p4[0] + v[0] * tc;
  Q[1] = p4[0] + v[1] * tc;
  Q[2] = p4[0] + v[2] * tc;
  double dP[3];
  for (int k = 0; k < 3; ++k) dP[k] = w[k] + (sc * u[k]) - (tc * v[k]);
  return sqrt(OR_DOT(dP, dP));
}

/* zoltan/hybrid.cpp:447-644 */
static double zh_pt(const double T1[3], const double T2[3], const double T3[3],
                    const double cp[3], double tq[3]) {
  double E0[3], E1[3], D[3], s, t;
  sub3(T2, T1, E0);
  sub3(T3, T1, E1);
  sub3(T1, cp, D);
  double sqr = pt_tri_sqr(OR_DOT(E0, E0), OR_DOT(E0, E1), OR_DOT(E1, E1), OR_DOT(E0, D),
                          OR_DOT(E1, D), OR_DOT(D, D), &s, &t);
  for (int k = 0; k < 3; ++k) tq[k] = T1[k] + (E1[k] * t) + (E0[k] * s);
  return sqrt(sqr);
}

/* zoltan/hybrid.cpp:646-885; P and Q in/out */
static void zh_bf(const double A[3], const double B[3], const double C[3], const double D[3],
                  const double E[3], const double F[3], double P[3], double Q[3]) {
  int hit = zh_segt(D, E, A, B, C, P) == 1;
  if (!hit) hit = zh_segt(E, F, A, B, C, P) == 1;
  if (!hit) hit = zh_segt(F, D, A, B, C, P) == 1;
  if (!hit) hit = zh_segt(A, B, D, E, F, P) == 1;
  if (!hit) hit = zh_segt(B, C, D, E, F, P) == 1;
  if (!hit) hit = zh_segt(C, A, D, E, F, P) == 1;
  if (hit) {
    for (int k = 0; k < 3; ++k) Q[k] = P[k];
    return;
  }
  const double* ta[3] = {A, B, C};
  const double* tb[3] = {D, E, F};
  double tp[3] = {0, 0, 0}, tq[3] = {0, 0, 0}, tmpP[3], tmpQ[3], ssP[3] = {0, 0, 0},
         ssQ[3] = {0, 0, 0};
  double ptmin = zh_pt(D, E, F, A, tmpP);
  for (int k = 0; k < 3; ++k) { tp[k] = A[k]; tq[k] = tmpP[k]; }
  for (int v = 1; v < 3; ++v) {
    const double tmp = zh_pt(D, E, F, ta[v], tmpP);
    if (ptmin > tmp) {
      ptmin = tmp;
      for (int k = 0; k < 3; ++k) { tp[k] = ta[v][k]; tq[k] = tmpP[k]; }
    }
  }
  for (int v = 0; v < 3; ++v) {
    const double tmp = zh_pt(A, B, C, tb[v], tmpP);
    if (ptmin > tmp) {
      ptmin = tmp;
      for (int k = 0; k < 3; ++k) { tq[k] = tb[v][k]; tp[k] = tmpP[k]; }
    }
  }
  double ssmin = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double tmp = zh_segseg(ta[i], ta[(i + 1) % 3], tb[j], tb[(j + 1) % 3], tmpP, tmpQ);
      if ((i == 0 && j == 0) || ssmin > tmp) {
        ssmin = tmp;
        for (int k = 0; k < 3; ++k) { ssP[k] = tmpP[k]; ssQ[k] = tmpQ[k]; }
      }
    }
  if (ssmin < ptmin) {
    for (int k = 0; k < 3; ++k) { P[k] = ssP[k]; Q[k] = ssQ[k]; }
  } else if (ssmin > ptmin) {
    for (int k = 0; k < 3; ++k) { P[k] = tp[k]; Q[k] = tq[k]; }
  } else if (ssmin == ptmin) {
    for (int k = 0; k < 3; ++k) { P[k] = ssP[k]; Q[k] = tq[k]; }
  }
}

/* zoltan/hybrid.cpp:37-132, serial semantics */
void or_zoltan_hybrid(int it, int nt, const double a[3], const double b[3], const double c[3],
                      double* const d[3], double* const e[3], double* const f[3],
                      double* const p[3], double* const q[3]) {
  double A[3] = {a[0], a[1], a[2]}, B[3] = {b[0], b[1], b[2]}, C[3] = {c[0], c[1], c[2]};
  double D[3], E[3], F[3], P[3] = {0, 0, 0}, Q[3] = {0, 0, 0}, err[4];
  for (unsigned k = (unsigned)it; k < (unsigned)nt / 4u; ++k) {
    for (unsigned l = 0; l < 4; ++l) {
      const unsigned i = k * 4 + l;
      for (int m = 0; m < 3; ++m) { D[m] = d[m][i]; E[m] = e[m][i]; F[m] = f[m][i]; }
      zh_penalty(A, B, C, D, E, F, P, Q, &err[l]);
      for (int m = 0; m < 3; ++m) { p[m][i] = P[m]; q[m][i] = Q[m]; }
    }
    double batch = 0.0;
    for (int l = 0; l < 4; ++l) batch += err[l];
    if (batch / 4u > 1E-8) {
      for (unsigned l = 0; l < 4; ++l) {
        const unsigned i = k * 4 + l;
        zh_bf(A, B, C, D, E, F, P, Q);
        for (int m = 0; m < 3; ++m) { p[m][i] = P[m]; q[m][i] = Q[m]; }
      }
    }
  }
}

/* zoltan/contact.cpp:56-240 with triangle i's y and z read (the reference
 * leaves Ay, Az uninitialised) and early-returning bf pairs skipped.
 * t[v][axis][i]; two-range form when all == 0.  Returns the number of kept
 * contacts (written up to cap) in discovery order. */
int or_zoltan_detection(int s1, int e1, int s2, int e2, int all, double* const t[6][3],
                        const int* pid, dh_zoltan_contact* out, int cap) {
  int n = 0;
  for (int i = s1; i < e1; ++i) {
    double Ax[3], Ay[3], Az[3];
    for (int v = 0; v < 3; ++v) { Ax[v] = t[v][0][i]; Ay[v] = t[v][1][i]; Az[v] = t[v][2][i]; }
    const int jb = all ? i + 1 : s2, je = all ? e1 : e2;
    for (int j = jb; j < je; ++j) {
      if (pid[i] == pid[j]) continue;
      double Bx[3], By[3], Bz[3], P[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
      for (int v = 0; v < 3; ++v) { Bx[v] = t[v][0][j]; By[v] = t[v][1][j]; Bz[v] = t[v][2][j]; }
      if (or_zoltan_bf(Ax, Ay, Az, Bx, By, Bz, P, Q) < 0) continue;
      const double dist = sqrt((Q[0] - P[0]) * (Q[0] - P[0]) + (Q[1] - P[1]) * (Q[1] - P[1]) +
                               (Q[2] - P[2]) * (Q[2] - P[2]));
      const double margin = 15e-2;
      if (!(dist < margin)) continue;
      int found = 0;
      for (int o = 0; o < n && o < cap; ++o) {
        const dh_zoltan_contact* k = out + o;
        if (k->pid[0] != pid[i]) continue;
        if (k->pp[0] == P[0] || k->pp[1] == P[1] || k->pp[2] == P[2] || k->qq[0] == Q[0] ||
            k->qq[1] == Q[1] || k->qq[2] == Q[2])
          found = 1;
      }
      if (found) continue;
      if (n < cap) {
        dh_zoltan_contact* k = out + n;
        const double depth = margin - dist;
        k->pid[0] = pid[i]; k->pid[1] = pid[j];
        k->color[0] = k->color[1] = 0;
        for (int m = 0; m < 3; ++m) {
          k->point[m] = (P[m] + Q[m]) / 2;
          k->normal[m] = (Q[m] - P[m]) / depth;
          k->pp[m] = P[m];
          k->qq[m] = Q[m];
        }
        k->depth = depth;
      }
      ++n;
    }
  }
  return n;
}

/* ---------------------------------------------------------------------------
 * ISPC kernels (delta/contact/detection/bf.ispc:3-937, penalty.ispc:3-174),
 * REAL = double.  Their algo.h is not in the reference; DOT and SUB are taken
 * as the Solfec-family definitions the reference carries elsewhere
 * (zoltan/algo.h:90-96,165-166): DOT(a,b) = (a0*b0 + a1*b1 + a2*b2),
 * SUB(a,b,c): c = a - b.  Arrays are vertex-major SoA: pp[v][k][i] is
 * coordinate k of vertex v of pair i.
 * ------------------------------------------------------------------------- */

/* Segment p1-p2 (origin p2) against p3-p4 (origin p4), SMALL_NUM 1e-7
 * (bf.ispc:360-440); returns the distance of the clamped parameters. */
static double ispc_segseg(const double p1[3], const double p2[3], const double p3[3],
                          const double p4[3], double* sc_out, double* tc_out) {
  double u[3], v[3], w[3];
  sub3(p1, p2, u);
  sub3(p3, p4, v);
  sub3(p2, p4, w);
  const double a = OR_DOT(u, u), b = OR_DOT(u, v), c = OR_DOT(v, v), d = OR_DOT(u, w),
               e = OR_DOT(v, w);
  const double D = (a * c - b * b);
  double sD = D, tD = D, sN = 0, tN = 0;
  const double small = 0.0000001;
  if (D < small) {
    sN = 0.0; sD = 1.0; tN = e; tD = c;
  } else {
    sN = (b * e - c * d);
    tN = (a * e - b * d);
    if (sN < 0.0) { sN = 0.0; tN = e; tD = c; }
    else if (sN > sD) { sN = sD; tN = e + b; tD = c; }
  }
  if (tN < 0.0) {
    tN = 0.0;
    if (-d < 0.0) sN = 0.0;
    else if (-d > a) sN = sD;
    else { sN = -d; sD = a; }
  } else if (tN > tD) {
    tN = tD;
    if ((-d + b) < 0.0) sN = 0;
    else if ((-d + b) > a) sN = sD;
    else { sN = (-d + b); sD = a; }
  }
  const double sc = (fabs(sN) < small) ? 0.0 : sN / sD;
  const double tc = (fabs(tN) < small) ? 0.0 : tN / tD;
  double dP[3];
  for (int k = 0; k < 3; ++k) dP[k] = w[k] + (sc * u[k]) - (tc * v[k]);
  *sc_out = sc;
  *tc_out = tc;
  return sqrt(dP[0] * dP[0] + dP[1] * dP[1] + dP[2] * dP[2]);
}

/* bf.ispc:9-935 for one pair: v[0..2] triangle A (pp1..pp3), v[3..5] B.
 * P, Q and *distance are in/out: the early returns on a zero normal write
 * only distance = -1 (bf.ispc:65-68,110-113); when the edge-edge minimum is
 * not beaten (ptmin < ssmin fails, bf.ispc:874-877) distance keeps the code
 * of the plane pre-pass (2, -2, -1 or 0, bf.ispc:114-174) while P and Q hold
 * the edge-edge points.  The pre-pass reads nn[3][2], w0[3][6] and dir[3][6]
 * through DOT on a row, so it uses nn[0][0], nn[0][1], nn[1][0] (or
 * nn[1][0], nn[1][1], nn[2][0]) and the x components of columns 0..2 of w0
 * and dir for every j; u and v are A's edges there.  An edge-edge index
 * that is never set (every distance NaN or >= 1e30) selects no case. */
static void ispc_bf_one(const double v[6][3], double P[3], double Q[3], double* distance) {
  const double *p1 = v[0], *p2 = v[1], *p3 = v[2], *p4 = v[3], *p5 = v[4], *p6 = v[5];
  double u[3], vv[3], nn[6], w0[18], dir[18];
  sub3(p5, p4, u);
  sub3(p6, p5, vv);
  nn[0] = u[1] * vv[2] - u[2] * vv[1];
  nn[2] = u[2] * vv[0] - u[0] * vv[2];
  nn[4] = u[0] * vv[1] - u[1] * vv[0];
  if (nn[0] == 0 && nn[2] == 0 && nn[4] == 0) { *distance = -1; return; }
  const double* pa[6] = {p1, p2, p3, p4, p5, p6};
  /* columns 0..2: A's edges against B's plane; 3..5: B's edges against A's */
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) {
      dir[6 * k + j] = pa[(j + 1) % 3][k] - pa[j][k];
      w0[6 * k + j] = pa[j][k] - p4[k];
    }
  sub3(p2, p1, u);
  sub3(p3, p1, vv);
  nn[1] = u[1] * vv[2] - u[2] * vv[1];
  nn[3] = u[2] * vv[0] - u[0] * vv[2];
  nn[5] = u[0] * vv[1] - u[1] * vv[0];
  if (nn[0] == 0 && nn[3] == 0 && nn[5] == 0) { *distance = -1; return; }
  for (int j = 3; j < 6; ++j)
    for (int k = 0; k < 3; ++k) {
      dir[6 * k + j] = pa[3 + (j - 3 + 1) % 3][k] - pa[j][k];
      w0[6 * k + j] = pa[j][k] - p1[k];
    }
  for (int j = 0; j < 6; ++j) {
    const double* nr = (j < 3) ? nn : nn + 2;
    const double a = -(nr[0] * w0[0] + nr[1] * w0[1] + nr[2] * w0[2]);
    const double b = (nr[0] * dir[0] + nr[1] * dir[1] + nr[2] * dir[2]);
    if (fabs(b) < 1E-30 && a == 0) { *distance = 2; continue; }
    const double r = a / b;
    for (int k = 0; k < 3; ++k) {
      P[k] = pa[j][k] + r * dir[6 * k + j];
      Q[k] = P[k];
    }
    if (r < 0.0 || r > 1.0 || r == INFINITY) { *distance = -2; break; }
    double w[3];
    sub3(P, j < 3 ? p4 : p1, w);
    const double uv = OR_DOT(u, vv), uu = OR_DOT(u, u), vvv = OR_DOT(vv, vv);
    const double D = uv * uv - uu * vvv;
    const double s = (uv * OR_DOT(w, vv) - vvv * OR_DOT(w, u)) / D;
    if (s < 0.0 || s > 1.0) { *distance = -1; break; }
    const double t = (uv * OR_DOT(w, u) - uu * OR_DOT(w, vv)) / D;
    if (t < 0.0 || (s + t) > 1.0) { *distance = -1; break; }
    *distance = 0;
  }
  /* nine edge pairs: A's edge k (origin A[k+1]) x B's edge m (origin B[m+1]) */
  double ssmin = 1E+30, ttc = 0, ssc = 0;
  int ssid = -1;
  for (int j = 0; j < 9; ++j) {
    const int ka = j / 3, kb = j % 3;
    double sc, tc;
    const double dist = ispc_segseg(pa[ka], pa[(ka + 1) % 3], pa[3 + kb], pa[3 + (kb + 1) % 3],
                                    &sc, &tc);
    if (dist < ssmin) { ssmin = dist; ttc = tc; ssc = sc; ssid = j; }
  }
  /* case 1 falls through into case 2 (bf.ispc:505-521) */
  if (ssid >= 0) {
    const int ka = ssid / 3, kb = (ssid == 1) ? 2 : ssid % 3;
    double ue[3], ve[3];
    sub3(pa[ka], pa[(ka + 1) % 3], ue);
    sub3(pa[3 + kb], pa[3 + (kb + 1) % 3], ve);
    for (int k = 0; k < 3; ++k) {
      P[k] = pa[(ka + 1) % 3][k] + (ue[k] * ssc);
      Q[k] = pa[3 + (kb + 1) % 3][k] + (ve[k] * ttc);
    }
  }
  /* six vertex-face pairs: A's face against p4, p5, p6, then B's against p1..p3 */
  double ss = 0, tt = 0, ptmin = 1E30;
  int id = 0;
  for (int j = 0; j < 6; ++j) {
    const double* f0 = j < 3 ? p1 : p4;
    const double* f1 = j < 3 ? p2 : p5;
    const double* f2 = j < 3 ? p3 : p6;
    const double* q = pa[j < 3 ? 3 + j : j - 3];
    double e0[3], e1[3], dd[3], s, t;
    sub3(f1, f0, e0);
    sub3(f2, f0, e1);
    sub3(f0, q, dd);
    const double sqr = pt_tri_sqr(OR_DOT(e0, e0), OR_DOT(e0, e1), OR_DOT(e1, e1), OR_DOT(e0, dd),
                                  OR_DOT(e1, dd), OR_DOT(dd, dd), &s, &t);
    const double dist = sqrt(sqr);
    if (dist < ptmin) { ptmin = dist; ss = s; tt = t; id = j; }
  }
  if (!(ptmin < ssmin)) return;
  *distance = ptmin;
  /* case 1 falls through into case 2 (bf.ispc:891-899): Q = p6 */
  if (id < 3) {
    for (int k = 0; k < 3; ++k) {
      P[k] = p1[k] + ((p3[k] - p1[k]) * tt) + ((p2[k] - p1[k]) * ss);
      Q[k] = pa[3 + (id == 1 ? 2 : id)][k];
    }
  } else {
    for (int k = 0; k < 3; ++k) {
      P[k] = pa[id - 3][k];
      Q[k] = p4[k] + ((p6[k] - p4[k]) * tt) + ((p5[k] - p4[k]) * ss);
    }
  }
}

void or_ispc_bf(unsigned n, double* const pp[6][3], double* const P[3], double* const Q[3],
                double* distance) {
  for (unsigned i = 0; i < n; ++i) {
    double v[6][3], p[3], q[3];
    for (int a = 0; a < 6; ++a)
      for (int k = 0; k < 3; ++k) v[a][k] = pp[a][k][i];
    for (int k = 0; k < 3; ++k) { p[k] = P[k][i]; q[k] = Q[k][i]; }
    ispc_bf_one(v, p, q, &distance[i]);
    for (int k = 0; k < 3; ++k) { P[k][i] = p[k]; Q[k][i] = q[k]; }
  }
}

/* penalty.ispc:3-142: 5 Newton steps, eps 0.1, r = lambda 10^(ln lambda + 14);
 * the regularisation is multiplied by r before the fifth step only (the
 * Newton index decides, penalty.ispc:97). */
void or_ispc_pen(unsigned n, double* const pp[6][3], double* const P[3], double* const Q[3],
                 double* distance) {
  for (unsigned i = 0; i < n; ++i) {
    double xa[3], ya[3], za[3], xb[3], yb[3], zb[3];
    for (int v = 0; v < 3; ++v) {
      xa[v] = pp[v][0][i]; ya[v] = pp[v][1][i]; za[v] = pp[v][2][i];
      xb[v] = pp[3 + v][0][i]; yb[v] = pp[3 + v][1][i]; zb[v] = pp[3 + v][2][i];
    }
    pen_ctx cx;
    pen_setup(xa, ya, za, xb, yb, zb, &cx);
    const double trace = cx.h[0] + cx.h[5] + cx.h[10] + cx.h[15];
    double delta = trace * 0.1;
    const double lambda = sqrt(0.0125 * trace);
    cx.r = lambda * pow(10, log(lambda) + 14);
    double x[4] = {0.33, 0.33, 0.33, 0.33};
    for (int j = 0; j < 5; ++j) {
      double dx[4];
      delta = j < 4 ? delta : cx.r * delta;
      pen_step(&cx, x, delta, dx);
      for (int k = 0; k < 4; ++k) x[k] = x[k] - dx[k];
    }
    double p[3], q[3], d[3];
    pen_points(&cx, x, p, q);
    sub3(p, q, d);
    distance[i] = sqrt(OR_DOT(d, d));
    for (int k = 0; k < 3; ++k) { P[k][i] = p[k]; Q[k][i] = q[k]; }
  }
}

/* ---------------------------------------------------------------------------
 * Adaptive time step of the Peano runner: triggerParticleTooClose
 * (multiscale/dem/mappings/Collision.cpp:188-299) over the candidate pairs,
 * then the travel speed (MoveParticles.cpp:84-89), finishedTimeStep,
 * adaptiveTimeStep and clearAccumulatedData (State.cpp:26-35,137-199).
 * MaxVelocityApproach is the maximum of -vBA over approaching pairs and the
 * previous travel speed (see dh_adaptive_state).
 * ------------------------------------------------------------------------- */
void or_adaptive_pairs(const int* pairs, int npairs, const double* pos, const double* linvel,
                       const double* radius, const double* eps, dh_adaptive_state* s) {
  double close = INFINITY, approach = s->max_vel_travel;
  int sep = 0;
  for (int k = 0; k < npairs; ++k) {
    const int a = pairs[2 * k], b = pairs[2 * k + 1];
    double pdt = s->dt * s->increase;
    if (pdt >= s->max_dt) pdt = s->max_dt;
    double cA[3], cB[3], dAB[3], pdAB[3], vA[3], vB[3];
    for (int q = 0; q < 3; ++q) {
      vA[q] = linvel[3 * a + q];
      vB[q] = linvel[3 * b + q];
      cA[q] = pos[3 * a + q] + pdt * vA[q];
      cB[q] = pos[3 * b + q] + pdt * vB[q];
    }
    for (int q = 0; q < 3; ++q) {
      dAB[q] = pos[3 * b + q] - pos[3 * a + q];
      pdAB[q] = cB[q] - cA[q];
    }
    const double d = sqrt((dAB[0] * dAB[0]) + (dAB[1] * dAB[1]) + (dAB[2] * dAB[2]));
    const double pd = sqrt((pdAB[0] * pdAB[0]) + (pdAB[1] * pdAB[1]) + (pdAB[2] * pdAB[2]));
    const double vBA = ((vB[0] * dAB[0]) + (vB[1] * dAB[1]) + (vB[2] * dAB[2])) -
                       ((vA[0] * dAB[0]) + (vA[1] * dAB[1]) + (vA[2] * dAB[2]));
    const double pvBA = ((vB[0] * pdAB[0]) + (vB[1] * pdAB[1]) + (vB[2] * pdAB[2])) -
                        ((vA[0] * pdAB[0]) + (vA[1] * pdAB[1]) + (vA[2] * pdAB[2]));
    if (vBA > 0 && pvBA > 0) { sep = 1; continue; }
    if (-vBA > approach) approach = -vBA;
    const double rA = ((2.0 * radius[a] + eps[a] * 2) * 1.1) / 2.0;
    const double rB = ((2.0 * radius[b] + eps[b] * 2) * 1.1) / 2.0;
    const double pdistancePerStep = (pd - rA - rB) / pdt;
    if (-pvBA >= pdistancePerStep) {
      const double rrA = (2.0 * radius[a]) / 2.0, rrB = (2.0 * radius[b]) / 2.0;
      double remaining = (d - rrA - rrB - eps[a] - eps[b]);
      if (remaining < 0.0) remaining = remaining * -1.1;
      double localMaxdt = remaining * 0.82;
      if (localMaxdt < 0.0) localMaxdt = -1 * localMaxdt;
      /* informStateThatTwoParticlesAreClose: the smallest positive proposal */
      if (localMaxdt > 0.0 && localMaxdt < close) close = localMaxdt;
    }
  }
  s->close = close < INFINITY ? close : 0.0;
  s->max_vel_approach = approach;
  s->separate = sep;
}

void or_adaptive_finish(const double* linvel, const int* obstacle, int n, dh_adaptive_state* s) {
  double travel = 0.0;
  for (int i = 0; i < n; ++i) {
    if (obstacle[i]) continue;
    const double* u = linvel + 3 * i;
    const double v = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (v > travel) travel = v;
  }
  s->time = s->time + s->dt;
  const double inc = s->increase, mindt = s->min_dt;
  s->max_dt = s->mesh_width / (2.0 * inc * s->max_vel_approach) * 0.80;
  int clamp = 1;
  if (s->close > 0.0) {
    if (s->dt >= mindt) {
      if (!(s->close > inc * s->dt)) s->dt = s->close;
      clamp = 0;
    } else if (s->dt < mindt) {
      s->dt = mindt;
    }
  } else if (s->step > 2) {
    s->dt = inc * s->dt;
  }
  if (clamp && s->dt > s->max_dt) s->dt = s->max_dt;
  s->max_vel_travel = travel;
  s->step = s->step + 1;
}
