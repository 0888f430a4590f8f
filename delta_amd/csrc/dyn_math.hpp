// dyn_math.hpp — binary64 force and rigid-body update arithmetic, shared by
// the fused HIP kernels and the host check build.  Operation order is the
// reference's (forces.cpp:208-220,255-396; dynamics.cpp:26-242); compiled
// without FMA contraction the results are bit-identical to it.
#pragma once
#include "cr_sincos.hpp"
#include "exact_math.hpp"

namespace dh {

// Exponential map of a rotation vector (dynamics.cpp:26-102), including the
// reference's Omega2*Omega3 product in the (1,3) term.
DH_HD void expmap(double o1, double o2, double o3, double L[9]) {
  const double v0 = o1 * o1, v1 = o2 * o2, v2 = o3 * o3;
  const double angsq = v0 + v1 + v2;
  double sx, cx;
  if (angsq < 3.0461741978671E-02) {
    // sin(x)/x and (1-cos x)/x^2 as series in x^2 (coefficients 1/(2k+1)!
    // and 1/(2k+2)!, 16 significant digits as dynamics.cpp:41-64)
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
    // the reference's libm sin and cos; on the device the correctly rounded
    // double-double ones (cr_sincos.hpp), which equal the host's where the
    // device libm's can differ in the last bit
    const double ang = dsqrt(angsq);
#if defined(__HIP_DEVICE_COMPILE__)
    sx = cr_sincos(ang, false) / ang;
    cx = (1.0 - cr_sincos(ang, true)) / angsq;
#else
    sx = std::sin(ang) / ang;
    cx = (1.0 - std::cos(ang)) / angsq;
#endif
  }
  const double v01 = o1 * o2, v02 = o2 * o3, v12 = o2 * o3;
  const double s0 = sx * o1, s1 = sx * o2, s2 = sx * o3;
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

// updateRotationMatrix (dynamics.cpp:104-138): R <- R expmap(h W), w = R W.
DH_HD void update_rotation(double ang[3], const double W[3], double R[9], double h) {
  double DL[9], r0[9];
  expmap(h * W[0], h * W[1], h * W[2], DL);
  for (int i = 0; i < 9; ++i) r0[i] = R[i];
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row)
      R[3 * col + row] = r0[row] * DL[3 * col] + r0[3 + row] * DL[3 * col + 1] +
                         r0[6 + row] * DL[3 * col + 2];
  for (int row = 0; row < 3; ++row)
    ang[row] = R[row] * W[0] + R[3 + row] * W[1] + R[6 + row] * W[2];
}

// updateAngular (dynamics.cpp:140-213): Solfec-style half-step update of the
// referential angular velocity W under spatial torque t.
DH_HD void update_angular(double W[3], const double R[9], const double J[9], const double Ji[9],
                          const double t[3], double h) {
  const double half = 0.5 * h;
  double T[3], DL[9], A[3], B[3];
  for (int row = 0; row < 3; ++row)
    T[row] = R[3 * row] * t[0] + R[3 * row + 1] * t[1] + R[3 * row + 2] * t[2];
  expmap(-half * W[0], -half * W[1], -half * W[2], DL);
  for (int r = 0; r < 3; ++r) A[r] = J[r] * W[0] + J[3 + r] * W[1] + J[6 + r] * W[2];
  for (int r = 0; r < 3; ++r) B[r] = DL[r] * A[0] + DL[3 + r] * A[1] + DL[6 + r] * A[2];
  for (int r = 0; r < 3; ++r) B[r] = B[r] + half * T[r];
  for (int r = 0; r < 3; ++r) A[r] = Ji[r] * B[0] + Ji[3 + r] * B[1] + Ji[6 + r] * B[2];
  for (int r = 0; r < 3; ++r) B[r] = J[r] * A[0] + J[3 + r] * A[1] + J[6 + r] * A[2];
  T[0] -= A[1] * B[2] - A[2] * B[1];
  T[1] -= A[2] * B[0] - A[0] * B[2];
  T[2] -= A[0] * B[1] - A[1] * B[0];
  for (int r = 0; r < 3; ++r) T[r] *= h;
  for (int r = 0; r < 3; ++r) W[r] = W[r] + Ji[r] * T[0] + Ji[3 + r] * T[1] + Ji[6 + r] * T[2];
}

// updateVertices for one vertex replica (dynamics.cpp:215-242).
DH_HD void update_vertex(double& x, double& y, double& z, double rx, double ry, double rz,
                         const double R[9], const double c[3], const double C[3]) {
  const double d0 = rx - C[0], d1 = ry - C[1], d2 = rz - C[2];
  x = c[0] + (R[0] * d0 + R[3] * d1 + R[6] * d2);
  y = c[1] + (R[1] * d0 + R[4] * d1 + R[7] * d2);
  z = c[2] + (R[2] * d0 + R[5] * d1 + R[8] * d2);
}

// One contact of getContactsForces (forces.cpp:320-376), accumulated into
// force/torque of the particle whose centre is posA.
struct ForceConsts {
  double spring, damper, friction;
};

DH_HD void contact_force(const double x[3], const double n[3], double depth, bool fric,
                         const double posA[3], const double angA[3], const double linA[3],
                         double massA, const double posB[3], const double angB[3],
                         const double linB[3], double massB, const ForceConsts& k,
                         double force[3], double torque[3]) {
  double z[3] = {x[0] - posA[0], x[1] - posA[1], x[2] - posA[2]};
  double vi[3], vj[3], vij[3];
  vi[0] = angA[1] * z[2] - angA[2] * z[1] + linA[0];
  vi[1] = angA[2] * z[0] - angA[0] * z[2] + linA[1];
  vi[2] = angA[0] * z[1] - angA[1] * z[0] + linA[2];
  z[0] = x[0] - posB[0];
  z[1] = x[1] - posB[1];
  z[2] = x[2] - posB[2];
  vj[0] = angB[1] * z[2] - angB[2] * z[1] + linB[0];
  vj[1] = angB[2] * z[0] - angB[0] * z[2] + linB[1];
  vj[2] = angB[0] * z[1] - angB[1] * z[0] + linB[2];
  for (int q = 0; q < 3; ++q) vij[q] = vj[q] - vi[q];
  const double ma = 1.0 / ((1.0 / massA) + (1.0 / massB));
  const double vel = (vij[0] * n[0]) + (vij[1] * n[1]) + (vij[2] * n[2]);
  const double damp = k.damper * 2.0 * dsqrt(ma * k.spring) * vel;
  const double fn = k.spring * depth + damp;
  const double f[3] = {fn * n[0], fn * n[1], fn * n[2]};
  double fr[3] = {0.0, 0.0, 0.0};
  if (fric) {
    for (int q = 0; q < 3; ++q) {
      const double vt = vi[q] - n[q] * ((vi[0] * n[0]) + (vi[1] * n[1]) + (vi[2] * n[2]));
      fr[q] = -vt * k.friction * fn;
    }
  }
  for (int q = 0; q < 3; ++q) force[q] += f[q] + fr[q];
  const double arm[3] = {x[0] - posA[0], x[1] - posA[1], x[2] - posA[2]};
  torque[0] += arm[1] * (f[2]) - arm[2] * (f[1]);
  torque[1] += arm[2] * (f[0]) - arm[0] * (f[2]);
  torque[2] += arm[0] * (f[1]) - arm[1] * (f[0]);
}

}  // namespace dh
