// shim_test.cpp — calls the reference-signature shim (delta_shim.cpp, built
// against the reference's headers) the way Delta's engine does and writes
// every result as hex doubles to stdout; the GPU test compares them with the
// oracle.  One process runs a whole script of commands read from stdin, one
// command per record (a letter, then its numbers):
//   D method nA nB <9nA> <9nB>   a detection driver (method = the DH_* code of
//                                bf, penalty, hybrid*, *Stat or gjk); prints the
//                                contact count, one line per contact, then the
//                                statistics accessors' values
//   C                            cleanHybridStatistics + cleanPenaltyStatistics
//   B <18>                       bfSolver on xA yA zA xB yB zB (3 each)
//   P maxErr <18>                penaltySolver; prints P, Q and `failed`
//   A <3 refAngular> <9 rotation> <9 inertia> <9 inverse> <3 torque> step
//                                updateAngular; prints refAngular
//   R <3 refAngular> <9 rotation> step
//                                updateRotationMatrix; prints angular, rotation
//   V refx refy refz <9 rotation> <3 position> <3 refposition>
//                                updateVertices; prints x y z
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "contact/contactpoint.h"
#include "contact/detection/bf.h"
#include "contact/detection/gjk.h"
#include "contact/detection/hybrid.h"
#include "contact/detection/penalty.h"
#include "dynamics/dynamics.h"

namespace {

bool read(double* v, int n) {
  for (int i = 0; i < n; ++i)
    if (scanf("%lf", v + i) != 1) return false;
  return true;
}

void print(const double* v, int n) {
  for (int i = 0; i < n; ++i) printf(i ? " %a" : "%a", v[i]);
  printf("\n");
}

int detect(int method) {
  int nA, nB;
  if (scanf("%d %d", &nA, &nB) != 2) return 2;
  std::vector<double> a(9 * static_cast<size_t>(nA) + 1), b(9 * static_cast<size_t>(nB) + 1);
  if (!read(a.data(), 9 * nA) || !read(b.data(), 9 * nB)) return 2;
  double* xA = a.data(); double* yA = xA + 3 * nA; double* zA = yA + 3 * nA;
  double* xB = b.data(); double* yB = xB + 3 * nB; double* zB = yB + 3 * nB;
  namespace d = delta::contact::detection;
  std::vector<delta::contact::contactpoint> r;
  switch (method) {
    case 0: r = d::bf(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 1: r = d::penalty(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 2:
      r = d::hybridWithPerBatchFallBack(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4,
                                        true, 9);
      break;
    case 3:
      r = d::hybridWithPerTriangleFallBack(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4,
                                           true, 9);
      break;
    case 4: r = d::penaltyStat(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 5: r = d::hybridBatchStat(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 6:
      r = d::hybridTriangleStat(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9);
      break;
    case 7: r = d::gjk(nA, xA, yA, zA, 1e-4, true, 7, nB, xB, yB, zB, 1e-4, true, 9); break;
    default: return 3;
  }
  printf("%zu\n", r.size());
  for (const auto& c : r) {
    const double v[14] = {c.x[0], c.x[1], c.x[2], c.normal[0], c.normal[1], c.normal[2], c.depth,
                          c.P[0], c.P[1], c.P[2], c.Q[0], c.Q[1], c.Q[2], c.epsilonTotal};
    print(v, 14);
    printf("%d %d %d\n", c.master, c.slave, c.friction ? 1 : 0);
  }
  const std::vector<int> h = d::getPenaltyIterationStatistics();
  for (int v : h) printf("%d ", v);
  printf("%d %d %d %a\n", d::getPenaltyFails(), d::getBatchFails(), d::getBatchSize(),
         d::getBatchError());
  return 0;
}

}  // namespace

int main() {
  char cmd[2];
  while (scanf("%1s", cmd) == 1) {
    switch (cmd[0]) {
      case 'D': {
        int method;
        if (scanf("%d", &method) != 1) return 2;
        const int rc = detect(method);
        if (rc) return rc;
        break;
      }
      case 'C':
        delta::contact::detection::cleanHybridStatistics();
        delta::contact::detection::cleanPenaltyStatistics();
        break;
      case 'B':
      case 'P': {
        double maxErr = 0, t[18], out[6];
        bool failed = false;
        if (cmd[0] == 'P' && !read(&maxErr, 1)) return 2;
        if (!read(t, 18)) return 2;
        if (cmd[0] == 'B')
          delta::contact::detection::bfSolver(t, t + 3, t + 6, t + 9, t + 12, t + 15, out[0],
                                              out[1], out[2], out[3], out[4], out[5]);
        else
          delta::contact::detection::penaltySolver(t, t + 3, t + 6, t + 9, t + 12, t + 15,
                                                   out[0], out[1], out[2], out[3], out[4],
                                                   out[5], maxErr, failed);
        print(out, 6);
        if (cmd[0] == 'P') printf("%d\n", failed ? 1 : 0);
        break;
      }
      case 'A': {
        double W[3], R[9], J[9], Ji[9], T[3], h;
        if (!read(W, 3) || !read(R, 9) || !read(J, 9) || !read(Ji, 9) || !read(T, 3) ||
            !read(&h, 1))
          return 2;
        delta::dynamics::updateAngular(W, R, J, Ji, T, h);
        print(W, 3);
        break;
      }
      case 'R': {
        double W[3], R[9], w[3] = {0, 0, 0}, h;
        if (!read(W, 3) || !read(R, 9) || !read(&h, 1)) return 2;
        delta::dynamics::updateRotationMatrix(w, W, R, h);
        print(w, 3);
        print(R, 9);
        break;
      }
      case 'V': {
        double r[3], R[9], p[3], rp[3], v[3];
        if (!read(r, 3) || !read(R, 9) || !read(p, 3) || !read(rp, 3)) return 2;
        delta::dynamics::updateVertices(&v[0], &v[1], &v[2], &r[0], &r[1], &r[2], R, p, rp);
        print(v, 3);
        break;
      }
      default:
        return 3;
    }
  }
  return 0;
}
