/*
 * cpu_baseline.c — OpenMP drivers over the oracle for the CPU baseline timed
 * beside the GPU (test/benchmark infrastructure only).
 *
 * or_tri_pairs_omp     the unique-pair workload, one thread per block of pairs.
 * or_detect_rows_omp   the reference-shaped particle-pair driver: OpenMP
 *                      `parallel for` over the triangles of A inside one pair
 *                      (bf.cpp:92-95 built with -DOMPTriangle); the contact
 *                      append is serialised like the reference's `critical`.
 * or_step_rows_omp     a whole DEM step (or_step's order: all-pairs boxes,
 *                      brute-force detection, forces, update) with that
 *                      row-parallel detection: the CPU baseline of the DEM
 *                      workloads.  Contacts are appended in completion order,
 *                      as in the reference, then each pair's are sorted back
 *                      into row-major order, so the step equals or_step's
 *                      bit for bit (and can check the GPU's on scenes the
 *                      serial oracle would take minutes over).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
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

/* Contacts of one particle pair: rows of A in parallel, each row's contacts
 * appended under a critical section (bf.cpp:92-95,125-141). */
/* Row-major order (triA, then triB) of one pair's contacts: the serial
 * driver's order, so that the step after the parallel rows is the serial
 * step bit for bit whatever order the critical section saw. */
static int contact_order(const void* a, const void* b) {
  const dh_contact* x = (const dh_contact*)a;
  const dh_contact* y = (const dh_contact*)b;
  if (x->triA != y->triA) return x->triA < y->triA ? -1 : 1;
  return (x->triB > y->triB) - (x->triB < y->triB);
}

static int rows_detect(const double* xA, const double* yA, const double* zA, int nA, double epsA,
                       int idA, const double* xB, const double* yB, const double* zB, int nB,
                       double epsB, int idB, int fric, dh_contact** buf, int* cap) {
  int n = 0;
  const double margin = epsA + epsB;
#pragma omp parallel for schedule(static)
  for (int iA = 0; iA < nA; ++iA) {
    dh_contact local[8];
    int nl = 0;
    for (int iB = 0; iB < nB; ++iB) {
      double P[3], Q[3];
      or_bf_solver(xA + 3 * iA, yA + 3 * iA, zA + 3 * iA, xB + 3 * iB, yB + 3 * iB, zB + 3 * iB,
                   P, Q);
      const double dx = Q[0] - P[0], dy = Q[1] - P[1], dz = Q[2] - P[2];
      if (sqrt(dx * dx + dy * dy + dz * dz) > margin) continue;
      or_contact_make(P[0], P[1], P[2], epsA, idA, Q[0], Q[1], Q[2], epsB, idB, fric,
                      local + nl);
      local[nl].triA = iA;
      local[nl].triB = iB;
      if (++nl == 8 || iB == nB - 1) {
#pragma omp critical
        {
          if (n + nl > *cap) {
            *cap = 2 * (n + nl);
            *buf = (dh_contact*)realloc(*buf, sizeof(dh_contact) * (size_t)*cap);
          }
          memcpy(*buf + n, local, sizeof(dh_contact) * (size_t)nl);
          n += nl;
        }
        nl = 0;
      }
    }
    if (nl) {
#pragma omp critical
      {
        if (n + nl > *cap) {
          *cap = 2 * (n + nl);
          *buf = (dh_contact*)realloc(*buf, sizeof(dh_contact) * (size_t)*cap);
        }
        memcpy(*buf + n, local, sizeof(dh_contact) * (size_t)nl);
        n += nl;
      }
    }
  }
  qsort(*buf, (size_t)n, sizeof(dh_contact), contact_order);
  return n;
}

int64_t or_step_rows_omp(or_scene* s, const dh_force_params* fp, double dt,
                         const double gravity[3], int64_t* checks, int* npairs_out,
                         int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  const int n = s->n;
  double* lo = (double*)malloc(sizeof(double) * 3 * n);
  double* hi = (double*)malloc(sizeof(double) * 3 * n);
  for (int p = 0; p < n; ++p) {
    for (int q = 0; q < 3; ++q) {
      lo[3 * p + q] = INFINITY;
      hi[3 * p + q] = -INFINITY;
    }
    for (int v = 3 * s->tri_offset[p]; v < 3 * s->tri_offset[p + 1]; ++v) {
      const double c[3] = {s->x[v], s->y[v], s->z[v]};
      for (int q = 0; q < 3; ++q) {
        lo[3 * p + q] = fmin(lo[3 * p + q], c[q]);
        hi[3 * p + q] = fmax(hi[3 * p + q], c[q]);
      }
    }
    const double m = s->eps[p] + 0.5e-9;
    for (int q = 0; q < 3; ++q) {
      lo[3 * p + q] -= m;
      hi[3 * p + q] += m;
    }
  }
  double* force = (double*)calloc(3 * (size_t)n, sizeof(double));
  double* torque = (double*)calloc(3 * (size_t)n, sizeof(double));
  int* touched = (int*)calloc((size_t)n, sizeof(int));
  int cap = 4096;
  dh_contact* buf = (dh_contact*)malloc(sizeof(dh_contact) * (size_t)cap);
  int64_t total = 0;
  int npairs = 0;
  const dh_bodies* b = &s->b;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (b->obstacle[i] && b->obstacle[j]) continue;
      int ov = 1;
      for (int q = 0; q < 3; ++q)
        if (lo[3 * i + q] > hi[3 * j + q] || lo[3 * j + q] > hi[3 * i + q]) ov = 0;
      if (!ov) continue;
      npairs++;
      const int oi = s->tri_offset[i], oj = s->tri_offset[j];
      const int ni = s->tri_offset[i + 1] - oi, nj = s->tri_offset[j + 1] - oj;
      if (checks) *checks += (int64_t)ni * nj;
      const int nc = rows_detect(s->x + 3 * oi, s->y + 3 * oi, s->z + 3 * oi, ni, s->eps[i],
                                 s->id[i], s->x + 3 * oj, s->y + 3 * oj, s->z + 3 * oj, nj,
                                 s->eps[j], s->id[j], s->friction[i] && s->friction[j], &buf,
                                 &cap);
      if (nc == 0) continue;
      total += nc;
      double fA[3] = {0, 0, 0}, tA[3] = {0, 0, 0}, fB[3] = {0, 0, 0}, tB[3] = {0, 0, 0};
      or_contacts_forces(buf, nc, 0, b->pos + 3 * i, b->angvel + 3 * i, b->linvel + 3 * i,
                         b->mass[i], b->pos + 3 * j, b->angvel + 3 * j, b->linvel + 3 * j,
                         b->mass[j], fp, fA, tA);
      or_contacts_forces(buf, nc, 1, b->pos + 3 * j, b->angvel + 3 * j, b->linvel + 3 * j,
                         b->mass[j], b->pos + 3 * i, b->angvel + 3 * i, b->linvel + 3 * i,
                         b->mass[i], fp, fB, tB);
      for (int q = 0; q < 3; ++q) {
        force[3 * i + q] += fA[q];
        torque[3 * i + q] += tA[q];
        force[3 * j + q] += fB[q];
        torque[3 * j + q] += tB[q];
      }
      touched[i] = touched[j] = 1;
    }
  dh_bodies* bw = &s->b;
  for (int p = 0; p < n; ++p) {
    if (bw->obstacle[p] || !touched[p]) continue;
    for (int q = 0; q < 3; ++q) bw->linvel[3 * p + q] += dt * (force[3 * p + q] / bw->mass[p]);
    or_update_angular(bw->refangvel + 3 * p, bw->rot + 9 * p, bw->inertia + 9 * p,
                      bw->inverse + 9 * p, torque + 3 * p, dt);
  }
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    const int kinematic = bw->obstacle[p] != 0;
    if (kinematic) {
      const double* u = bw->linvel + 3 * p;
      const double* W = bw->refangvel + 3 * p;
      if (u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0 && W[0] == 0.0 && W[1] == 0.0 && W[2] == 0.0)
        continue;
    }
    for (int q = 0; q < 3; ++q)
      if (gravity[q] != 0.0 && !kinematic) bw->linvel[3 * p + q] += dt * gravity[q];
    for (int q = 0; q < 3; ++q) bw->pos[3 * p + q] += dt * bw->linvel[3 * p + q];
    or_update_rotation_matrix(bw->angvel + 3 * p, bw->refangvel + 3 * p, bw->rot + 9 * p, dt);
    for (int v = 3 * s->tri_offset[p]; v < 3 * s->tri_offset[p + 1]; ++v)
      or_update_vertex(s->x + v, s->y + v, s->z + v, s->rx[v], s->ry[v], s->rz[v],
                       bw->rot + 9 * p, bw->pos + 3 * p, bw->refpos + 3 * p);
  }
  free(lo); free(hi); free(force); free(torque); free(touched); free(buf);
  if (npairs_out) *npairs_out = npairs;
  return total;
}
