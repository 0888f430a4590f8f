// shim_test.cpp — calls the reference-signature shim (delta_shim.cpp, built
// against the reference's headers) the way Delta's engine does and writes the
// contacts as raw dh_contact-compatible doubles to stdout; the GPU test
// compares them with the oracle.  Input: nA nB then 9*nA and 9*nB doubles
// (x[3t+k] blocks of x, y, z) on stdin.
#include <cstdio>
#include <vector>

#include "contact/contactpoint.h"
#include "contact/detection/bf.h"
#include "contact/detection/gjk.h"
#include "contact/detection/hybrid.h"
#include "contact/detection/penalty.h"

int main(int argc, char** argv) {
  int method = argc > 1 ? atoi(argv[1]) : 0;
  int nA, nB;
  if (scanf("%d %d", &nA, &nB) != 2) return 2;
  std::vector<double> a(9 * nA), b(9 * nB);
  for (auto& v : a) if (scanf("%lf", &v) != 1) return 2;
  for (auto& v : b) if (scanf("%lf", &v) != 1) return 2;
  double* xA = a.data(); double* yA = xA + 3 * nA; double* zA = yA + 3 * nA;
  double* xB = b.data(); double* yB = xB + 3 * nB; double* zB = yB + 3 * nB;
  namespace d = delta::contact::detection;
  std::vector<delta::contact::contactpoint> r;
  switch (method) {
    case 0: r = d::bf(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 1: r = d::penalty(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 2: r = d::hybridWithPerBatchFallBack(xA, yA, zA, nA, 1e-4, true, 7, xB, yB, zB, nB, 1e-4, true, 9); break;
    case 7: r = d::gjk(nA, xA, yA, zA, 1e-4, true, 7, nB, xB, yB, zB, 1e-4, true, 9); break;
    default: return 3;
  }
  printf("%zu\n", r.size());
  for (const auto& c : r)
    printf("%a %a %a %a %a %a %a %a %a %a %a %a %a %a\n", c.x[0], c.x[1], c.x[2], c.normal[0],
           c.normal[1], c.normal[2], c.depth, c.P[0], c.P[1], c.P[2], c.Q[0], c.Q[1], c.Q[2],
           c.epsilonTotal);
  return 0;
}
