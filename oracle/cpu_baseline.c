/*
 * cpu_baseline.c — OpenMP drivers over the oracle for the CPU baseline timed
 * beside the GPU (test/benchmark infrastructure only).
 *
 * or_tri_pairs_omp     the unique-pair workload, one thread per block of pairs.
 * or_detect_rows_omp   the reference-shaped particle-pair driver: OpenMP
 *                      `parallel for` over the triangles of A inside one pair
 *                      (bf.cpp:92-95 built with -DOMPTriangle); the contact
 *                      append is serialised like the reference's `critical`.
 */
#include <omp.h>
#include <stdlib.h>
#include "oracle.h"

int or_tri_pairs_omp(int method, const double* soa, int64_t n, double eps_total,
                     double* P, double* Q, double* d, uint8_t* flags, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  const int64_t block = 4096;
  const int64_t nblocks = (n + block - 1) / block;
  int rc = DH_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t b = 0; b < nblocks; ++b) {
    int64_t i0 = b * block;
    int64_t m = (n - i0 < block) ? (n - i0) : block;
    /* gather the block into a private SoA so or_tri_pairs sees stride m */
    double* s = (double*)malloc(sizeof(double) * 18 * m);
    double* p = (double*)malloc(sizeof(double) * 3 * m);
    double* q = (double*)malloc(sizeof(double) * 3 * m);
    for (int k = 0; k < 18; ++k)
      for (int64_t i = 0; i < m; ++i) s[k * m + i] = soa[k * n + i0 + i];
    int r = or_tri_pairs(method, s, m, eps_total, p, q, d + i0, flags ? flags + i0 : NULL);
    if (r != DH_OK) rc = r;
    for (int c = 0; c < 3; ++c)
      for (int64_t i = 0; i < m; ++i) {
        P[c * n + i0 + i] = p[c * m + i];
        Q[c * n + i0 + i] = q[c * m + i];
      }
    free(s); free(p); free(q);
  }
  return rc;
}

/* Reference-shaped BF over one particle pair: rows of A in parallel, contact
 * count per row gathered in a critical section.  Returns the contact count. */
int64_t or_detect_rows_omp(const double* xA, const double* yA, const double* zA, int nA,
                           const double* xB, const double* yB, const double* zB, int nB,
                           double margin, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  int64_t total = 0;
#pragma omp parallel for schedule(static)
  for (int iA = 0; iA < nA; ++iA) {
    int64_t local = 0;
    for (int iB = 0; iB < nB; ++iB) {
      double P[3], Q[3];
      or_bf_solver(xA + 3 * iA, yA + 3 * iA, zA + 3 * iA, xB + 3 * iB, yB + 3 * iB, zB + 3 * iB,
                   P, Q);
      double dx = Q[0] - P[0], dy = Q[1] - P[1], dz = Q[2] - P[2];
      if (__builtin_sqrt(dx * dx + dy * dy + dz * dz) <= margin) local++;
    }
#pragma omp critical
    total += local;
  }
  return total;
}
