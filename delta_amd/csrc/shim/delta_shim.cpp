// delta_shim.cpp — drop-in replacement for the reference's
//   delta/contact/detection/{bf,penalty,hybrid}.cpp   (detection drivers and
//                                                      pair solvers)
//   delta/contact/detection/gjk.cpp                   (GJK wrapper and core)
//   delta/dynamics/dynamics.cpp                       (rigid-body update)
// A Delta build compiles this file instead of those five, against Delta's own
// headers (bf.h, penalty.h, hybrid.h, dynamics.h, contactpoint.h), and links
// libdelta_hip.so.  (forces.cpp stays: per call it is host arithmetic over a
// few contacts; the batched GPU form is dh_contacts_forces inside dh_step.)  Every function keeps the reference signature
// and meaning; the work runs in the HIP library through include/delta_hip.h.
// Statistics keep the reference's global-state semantics (not thread-safe;
// cleanHybridStatistics leaves the batch size and error, hybrid.cpp:36-41).
#include <cfloat>
#include <stdexcept>
#include <string>
#include <vector>

#include "contact/contactpoint.h"
#include "contact/detection/bf.h"
#include "contact/detection/gjk.h"
#include "contact/detection/hybrid.h"
#include "contact/detection/penalty.h"
#include "dynamics/dynamics.h"
#include "delta_hip.h"
#include "gjk.hpp"

namespace {

dh_stats g_stats = {};

delta::contact::contactpoint to_contactpoint(const dh_contact& c) {
  delta::contact::contactpoint r;
  for (int q = 0; q < 3; ++q) {
    r.x[q] = c.x[q];
    r.normal[q] = c.normal[q];
    r.P[q] = c.P[q];
    r.Q[q] = c.Q[q];
  }
  r.depth = c.depth;
  r.epsilonTotal = c.eps_total;
  r.friction = c.friction != 0;
  r.master = c.master;
  r.slave = c.slave;
  r.penetration = c.penetration;
  return r;
}

void check(int rc, const char* what) {
  if (rc != DH_OK) throw std::runtime_error(std::string(what) + ": " + dh_last_error());
}

std::vector<delta::contact::contactpoint> run(int method, const iREAL* xA, const iREAL* yA,
                                              const iREAL* zA, int nA, iREAL epsA, bool fA,
                                              int idA, const iREAL* xB, const iREAL* yB,
                                              const iREAL* zB, int nB, iREAL epsB, bool fB,
                                              int idB) {
  std::vector<dh_contact> buf(static_cast<size_t>(nA > 0 && nB > 0 ? 64 : 1));
  int n = 0;
  for (;;) {
    dh_stats saved = g_stats;
    const int rc = dh_detect_pair(method, xA, yA, zA, nA, epsA, fA, idA, xB, yB, zB, nB, epsB,
                                  fB, idB, buf.data(), static_cast<int>(buf.size()), &n, &g_stats);
    if (rc == DH_ERR_CAPACITY) {
      g_stats = saved;
      buf.resize(static_cast<size_t>(n));
      continue;
    }
    check(rc, "dh_detect_pair");
    break;
  }
  std::vector<delta::contact::contactpoint> out;
  out.reserve(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) out.push_back(to_contactpoint(buf[static_cast<size_t>(k)]));
  return out;
}

}  // namespace

namespace delta {
namespace contact {
namespace detection {

// gjk.cpp:25-100, on the GPU.  Uses every vertex of both particles: the
// reference copies nTri vertices and then reads 3*nTri points from that
// buffer, past its end.
std::vector<contactpoint> gjk(int nA, iREAL* xA, iREAL* yA, iREAL* zA, iREAL epsA, bool fA,
                              int pA, int nB, iREAL* xB, iREAL* yB, iREAL* zB, iREAL epsB,
                              bool fB, int pB) {
  return run(DH_GJK, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

// gjk.cpp:418-460 on two AoS point sets, on the host: the kernel's own
// iteration (gjk.hpp) with sequential support searches, so results equal
// the GPU wrapper's for the same points.
iREAL gjk(iREAL* a, int na, iREAL* b, int nb, iREAL* p, iREAL* q) {
  auto pa = [&](int i) { return dh::V3<double>{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; };
  auto pb = [&](int i) { return dh::V3<double>{b[3 * i], b[3 * i + 1], b[3 * i + 2]}; };
  auto search = [](const double* c, int n, const double* v, bool maximal) {
    int best = -1;
    double bd = maximal ? -DBL_MAX : DBL_MAX;
    for (int i = 0; i < n; ++i) {
      const double d = dh::gdot(c + 3 * i, v);
      if (maximal ? d > bd : d < bd) {
        bd = d;
        best = i;
      }
    }
    return best;
  };
  auto min_a = [&](const double* v) { return search(a, na, v, false); };
  auto max_b = [&](const double* v) { return search(b, nb, v, true); };
  bool ok = true;
  const double d = dh::gjk_core(na, nb, pa, pb, min_a, max_b, p, q, &ok);
  if (!ok) throw std::runtime_error("gjk: no support point (NaN coordinates)");
  return d;
}

std::vector<contactpoint> bf(const iREAL* xA, const iREAL* yA, const iREAL* zA, int nA,
                             iREAL epsA, bool fA, int pA, const iREAL* xB, const iREAL* yB,
                             const iREAL* zB, int nB, iREAL epsB, bool fB, int pB) {
  return run(DH_BF, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> penalty(const iREAL* xA, const iREAL* yA, const iREAL* zA, const int nA,
                                  const iREAL epsA, const bool fA, const int pA, const iREAL* xB,
                                  const iREAL* yB, const iREAL* zB, const int nB,
                                  const iREAL epsB, const bool fB, const int pB) {
  return run(DH_PENALTY, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> penaltyStat(const iREAL* xA, const iREAL* yA, const iREAL* zA,
                                      const int nA, const iREAL epsA, const bool fA, const int pA,
                                      const iREAL* xB, const iREAL* yB, const iREAL* zB,
                                      const int nB, const iREAL epsB, const bool fB,
                                      const int pB) {
  return run(DH_PENALTY_STAT, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> hybridWithPerBatchFallBack(const iREAL* xA, const iREAL* yA,
                                                     const iREAL* zA, const int nA,
                                                     const iREAL epsA, const bool fA,
                                                     const int pA, const iREAL* xB,
                                                     const iREAL* yB, const iREAL* zB,
                                                     const int nB, const iREAL epsB,
                                                     const bool fB, const int pB) {
  return run(DH_HYBRID_BATCH, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> hybridWithPerTriangleFallBack(const iREAL* xA, const iREAL* yA,
                                                        const iREAL* zA, const int nA,
                                                        const iREAL epsA, const bool fA,
                                                        const int pA, const iREAL* xB,
                                                        const iREAL* yB, const iREAL* zB,
                                                        const int nB, const iREAL epsB,
                                                        const bool fB, const int pB) {
  return run(DH_HYBRID_TRIANGLE, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> hybridBatchStat(const iREAL* xA, const iREAL* yA, const iREAL* zA,
                                          const int nA, const iREAL epsA, const bool fA,
                                          const int pA, const iREAL* xB, const iREAL* yB,
                                          const iREAL* zB, const int nB, const iREAL epsB,
                                          const bool fB, const int pB) {
  return run(DH_HYBRID_BATCH_STAT, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB, pB);
}

std::vector<contactpoint> hybridTriangleStat(const iREAL* xA, const iREAL* yA, const iREAL* zA,
                                             const int nA, const iREAL epsA, const bool fA,
                                             const int pA, const iREAL* xB, const iREAL* yB,
                                             const iREAL* zB, const int nB, const iREAL epsB,
                                             const bool fB, const int pB) {
  return run(DH_HYBRID_TRIANGLE_STAT, xA, yA, zA, nA, epsA, fA, pA, xB, yB, zB, nB, epsB, fB,
             pB);
}

// single-pair solvers: one-pair batches of the unique-pair kernels
static void one_pair(int method, const iREAL* xA, const iREAL* yA, const iREAL* zA,
                     const iREAL* xB, const iREAL* yB, const iREAL* zB, double eps_total,
                     iREAL& xPA, iREAL& yPA, iREAL& zPA, iREAL& xPB, iREAL& yPB, iREAL& zPB,
                     unsigned char* flag) {
  double soa[18], P[3], Q[3], d[1];
  for (int k = 0; k < 3; ++k) {
    soa[k] = xA[k];
    soa[3 + k] = yA[k];
    soa[6 + k] = zA[k];
    soa[9 + k] = xB[k];
    soa[12 + k] = yB[k];
    soa[15 + k] = zB[k];
  }
  void* buf = nullptr;
  check(dh_malloc(&buf, sizeof(double) * 25 + 8), "dh_malloc");
  double* dv = static_cast<double*>(buf);
  check(dh_memcpy_h2d(dv, soa, sizeof(soa)), "h2d");
  check(dh_tri_pairs(method, dv, nullptr, 1, eps_total, dv + 18, dv + 21, dv + 24,
                     reinterpret_cast<uint8_t*>(dv + 25), nullptr),
        "dh_tri_pairs");
  check(dh_memcpy_d2h(P, dv + 18, sizeof(P)), "d2h");
  check(dh_memcpy_d2h(Q, dv + 21, sizeof(Q)), "d2h");
  check(dh_memcpy_d2h(d, dv + 24, sizeof(d)), "d2h");
  if (flag) check(dh_memcpy_d2h(flag, dv + 25, 1), "d2h");
  dh_free(buf);
  xPA = P[0]; yPA = P[1]; zPA = P[2];
  xPB = Q[0]; yPB = Q[1]; zPB = Q[2];
}

void bfSolver(const iREAL* xA, const iREAL* yA, const iREAL* zA, const iREAL* xB,
              const iREAL* yB, const iREAL* zB, iREAL& xPA, iREAL& yPA, iREAL& zPA, iREAL& xPB,
              iREAL& yPB, iREAL& zPB) {
  one_pair(DH_PAIR_BF64, xA, yA, zA, xB, yB, zB, 0.0, xPA, yPA, zPA, xPB, yPB, zPB, nullptr);
}

void penaltySolver(const iREAL* xA, const iREAL* yA, const iREAL* zA, const iREAL* xB,
                   const iREAL* yB, const iREAL* zB, iREAL& xPA, iREAL& yPA, iREAL& zPA,
                   iREAL& xPB, iREAL& yPB, iREAL& zPB, iREAL maxErr, bool& failed) {
  unsigned char f = 0;
  one_pair(DH_PAIR_PENALTY64, xA, yA, zA, xB, yB, zB, 16.0 * maxErr, xPA, yPA, zPA, xPB, yPB,
           zPB, &f);
  failed = (f & DH_FLAG_FAILED) != 0;
}

void cleanPenaltyStatistics() {
  for (int i = 0; i < 17; ++i) g_stats.newton_histogram[i] = 0;
}

std::vector<int> getPenaltyIterationStatistics() {
  std::vector<int> h(17);
  for (int i = 0; i < 17; ++i) h[i] = static_cast<int>(g_stats.newton_histogram[i]);
  return h;
}

std::vector<std::vector<iREAL>> getPenaltyErrorStatistics() { return {}; }

void cleanHybridStatistics() {
  g_stats.penalty_fails = 0;
  g_stats.batch_fails = 0;
}
int getPenaltyFails() { return static_cast<int>(g_stats.penalty_fails); }
int getBatchFails() { return static_cast<int>(g_stats.batch_fails); }
int getBatchSize() { return static_cast<int>(g_stats.batch_size); }
iREAL getBatchError() { return g_stats.batch_error; }

}  // namespace detection
}  // namespace contact

namespace dynamics {

void updateAngular(iREAL* refAngular, iREAL* rotation, iREAL* inertia, iREAL* inverse,
                   iREAL* torque, iREAL step) {
  check(dh_update_angular(refAngular, rotation, inertia, inverse, torque, step),
        "dh_update_angular");
}

void updateRotationMatrix(iREAL* angular, iREAL* refAngular, iREAL* rotation, iREAL step) {
  check(dh_update_rotation_matrix(angular, refAngular, rotation, step),
        "dh_update_rotation_matrix");
}

// One vertex replica: kept on the host (a kernel launch per vertex would
// cost more than the arithmetic); the batched form is dh_integrate.
void updateVertices(iREAL* x, iREAL* y, iREAL* z, iREAL* refx, iREAL* refy, iREAL* refz,
                    iREAL* rotation, iREAL* position, iREAL* refposition) {
  const iREAL C0 = *refx - refposition[0], C1 = *refy - refposition[1],
              C2 = *refz - refposition[2];
  *x = position[0] + (rotation[0] * C0 + rotation[3] * C1 + rotation[6] * C2);
  *y = position[1] + (rotation[1] * C0 + rotation[4] * C1 + rotation[7] * C2);
  *z = position[2] + (rotation[2] * C0 + rotation[5] * C1 + rotation[8] * C2);
}

}  // namespace dynamics
}  // namespace delta
