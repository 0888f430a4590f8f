// ref_wrap.cpp — C entry points over the reference's own contact, force and
// dynamics functions, compiled from the sources where they lie under the
// reference tree (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: the result
// (oracle/_ref/libdelta_ref.so) pins the C restatement in oracle.c; it is never
// part of the product.  No reference source is copied here; this file only
// calls the reference's public functions and converts their std::vector
// results into dh_contact records.
#include <cstring>
#include <vector>

#include "contact/contactpoint.h"
#include "contact/detection/bf.h"
#include "contact/detection/gjk.h"
#include "contact/detection/hybrid.h"
#include "contact/detection/penalty.h"
#include "contact/forces/forces.h"
#include "dynamics/dynamics.h"

#include "../include/delta_hip.h"

namespace det = delta::contact::detection;
using delta::contact::contactpoint;

static void to_dh(const contactpoint& c, dh_contact* o) {
  std::memset(o, 0, sizeof(*o));
  for (int q = 0; q < 3; ++q) {
    o->x[q] = c.x[q];
    o->normal[q] = c.normal[q];
    o->P[q] = c.P[q];
    o->Q[q] = c.Q[q];
  }
  o->depth = c.depth;
  o->eps_total = c.epsilonTotal;
  o->master = c.master;
  o->slave = c.slave;
  o->penetration = c.penetration;
  o->friction = c.friction ? 1 : 0;
  o->triA = -1;
  o->triB = -1;
}

static contactpoint from_dh(const dh_contact& c) {
  contactpoint r;
  for (int q = 0; q < 3; ++q) {
    r.x[q] = c.x[q];
    r.normal[q] = c.normal[q];
    r.P[q] = c.P[q];
    r.Q[q] = c.Q[q];
  }
  r.depth = c.depth;
  r.epsilonTotal = c.eps_total;
  r.master = c.master;
  r.slave = c.slave;
  r.penetration = c.penetration;
  r.friction = c.friction != 0;
  return r;
}

extern "C" {

void ref_bf_solver(const double* xA, const double* yA, const double* zA, const double* xB,
                   const double* yB, const double* zB, double* P, double* Q) {
  det::bfSolver(xA, yA, zA, xB, yB, zB, P[0], P[1], P[2], Q[0], Q[1], Q[2]);
}

void ref_penalty_solver(const double* xA, const double* yA, const double* zA, const double* xB,
                        const double* yB, const double* zB, double* P, double* Q, double max_err,
                        int* failed) {
  bool f = false;
  det::penaltySolver(xA, yA, zA, xB, yB, zB, P[0], P[1], P[2], Q[0], Q[1], Q[2], max_err, f);
  *failed = f ? 1 : 0;
}

void ref_penalty_solver_adaptive(const double* xA, const double* yA, const double* zA,
                                 const double* xB, const double* yB, const double* zB, double* P,
                                 double* Q, double max_err, int* its) {
  int n = *its;
  det::penaltySolver(xA, yA, zA, xB, yB, zB, P[0], P[1], P[2], Q[0], Q[1], Q[2], max_err, n);
  *its = n;
}

// Runs one reference driver; returns the number of contacts (all of them are
// copied when cap allows).  stats receives the reference's global counters
// after the call (the caller cleans them with ref_clean_stats).
int ref_detect(int method, const double* xA, const double* yA, const double* zA, int nA,
               double epsA, int fricA, int idA, const double* xB, const double* yB,
               const double* zB, int nB, double epsB, int fricB, int idB, dh_contact* out,
               int cap) {
  std::vector<contactpoint> r;
  switch (method) {
    case DH_BF:
      r = det::bf(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB, epsB, fricB != 0, idB);
      break;
    case DH_PENALTY:
      r = det::penalty(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB, epsB, fricB != 0,
                       idB);
      break;
    case DH_HYBRID_BATCH:
      r = det::hybridWithPerBatchFallBack(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB,
                                          epsB, fricB != 0, idB);
      break;
    case DH_HYBRID_TRIANGLE:
      r = det::hybridWithPerTriangleFallBack(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB,
                                             nB, epsB, fricB != 0, idB);
      break;
    case DH_PENALTY_STAT:
      r = det::penaltyStat(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB, epsB,
                           fricB != 0, idB);
      break;
    case DH_HYBRID_BATCH_STAT:
      r = det::hybridBatchStat(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB, epsB,
                               fricB != 0, idB);
      break;
    case DH_HYBRID_TRIANGLE_STAT:
      r = det::hybridTriangleStat(xA, yA, zA, nA, epsA, fricA != 0, idA, xB, yB, zB, nB, epsB,
                                  fricB != 0, idB);
      break;
    case DH_GJK: {
      std::vector<double> xa(xA, xA + 3 * nA), ya(yA, yA + 3 * nA), za(zA, zA + 3 * nA);
      std::vector<double> xb(xB, xB + 3 * nB), yb(yB, yB + 3 * nB), zb(zB, zB + 3 * nB);
      r = det::gjk(nA, xa.data(), ya.data(), za.data(), epsA, fricA != 0, idA, nB, xb.data(),
                   yb.data(), zb.data(), epsB, fricB != 0, idB);
    } break;
    default:
      return -1;
  }
  for (size_t k = 0; k < r.size() && (int)k < cap; ++k) to_dh(r[k], out + k);
  return (int)r.size();
}

// The reference's GJK core on AoS point sets (gjk.cpp:418-460).  Its
// detection wrapper is not exposed: it passes 3*nTri as the point count of a
// buffer holding nTri points (gjk.cpp:58-77) and so reads past its allocation.
double ref_gjk_core(const double* a, int na, const double* b, int nb, double* p, double* q) {
  std::vector<double> ca(a, a + 3 * na), cb(b, b + 3 * nb);
  return det::gjk(ca.data(), na, cb.data(), nb, p, q);
}

void ref_clean_stats(void) {
  det::cleanPenaltyStatistics();
  det::cleanHybridStatistics();
}

void ref_get_stats(long long* histogram17, long long* penalty_fails, long long* batch_fails,
                   long long* batch_size, double* batch_error) {
  std::vector<int> h = det::getPenaltyIterationStatistics();
  for (int i = 0; i < 17; ++i) histogram17[i] = (i < (int)h.size()) ? h[i] : 0;
  *penalty_fails = det::getPenaltyFails();
  *batch_fails = det::getBatchFails();
  *batch_size = det::getBatchSize();
  *batch_error = det::getBatchError();
}

void ref_contacts_forces(const dh_contact* cs, int nc, int negate_normal, double* posA,
                         double* refposA, double* angA, double* refangA, double* linA,
                         double massA, double* invA, double* rotA, double* posB, double* refposB,
                         double* angB, double* refangB, double* linB, double massB, double* invB,
                         double* rotB, int is_sphere, double* force, double* torque) {
  std::vector<contactpoint> v;
  for (int k = 0; k < nc; ++k) {
    contactpoint c = from_dh(cs[k]);
    if (negate_normal)
      for (int q = 0; q < 3; ++q) c.normal[q] = -1.0 * c.normal[q];
    v.push_back(c);
  }
  delta::contact::forces::getContactsForces(v, posA, refposA, angA, refangA, linA, massA, invA,
                                            rotA, 0, posB, refposB, angB, refangB, linB, massB,
                                            invB, rotB, 0, force, torque, is_sphere != 0);
}

void ref_update_angular(double* refAngular, double* rotation, double* inertia, double* inverse,
                        double* torque, double step) {
  delta::dynamics::updateAngular(refAngular, rotation, inertia, inverse, torque, step);
}

void ref_update_rotation_matrix(double* angular, double* refAngular, double* rotation,
                                double step) {
  delta::dynamics::updateRotationMatrix(angular, refAngular, rotation, step);
}

void ref_update_vertices(double* x, double* y, double* z, double* rx, double* ry, double* rz,
                         double* rotation, double* position, double* refposition) {
  delta::dynamics::updateVertices(x, y, z, rx, ry, rz, rotation, position, refposition);
}

}  // extern "C"
