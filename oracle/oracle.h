/*
 * oracle.h — CPU restatement of the reference's contact-detection, force and
 * dynamics arithmetic.  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * HIP path is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it; the product never calls it.
 *
 * Every function is IEEE binary64, compiled without FMA contraction, in the
 * reference's operation order, so that results are bit-identical to the
 * reference built with g++ -O3 (no -march, hence no fused multiply-add;
 * delta/Makefile:22-23,41).
 *
 * Parity pins (how each function is checked):
 *   or_bf_solver            golden vectors from the compiled reference
 *                           (oracle/_ref), closed forms, symmetry, brute-force
 *                           minimum over the 15 feature pairs.
 *   or_penalty_solver*      golden vectors from oracle/_ref; convergence to the
 *                           brute-force optimum for well-separated pairs.
 *   or_detect (all methods) golden vectors from oracle/_ref.
 *   or_contact_make         golden vectors (contact fields) from oracle/_ref.
 *   or_contacts_forces      oracle/_ref + two-particle head-on closed form.
 *   or_expmap & dynamics    oracle/_ref + Rodrigues closed form (detects the
 *                           v02 term) + torque-free invariants.
 *   or_gjk                  oracle/_ref + equality with bfSolver minimum over
 *                           hull surfaces for separated convex sets.
 *   or_zoltan_bf            zoltan reference golden vectors (oracle/_ref) +
 *                           equality with or_bf_solver away from the
 *                           fall-through cases.
 *   or_zoltan_penalty/hybrid zoltan reference golden vectors (to libm rounding
 *                           of pow/log/log10) + convergence to the brute force.
 *   or_zoltan_detection     composition of or_zoltan_bf with the contact
 *                           formulas; hand-built de-duplication cases.
 *   or_ispc_bf              no ispc compiler here: the point-triangle solve is
 *                           or_bf_solver's (golden-pinned); the result equals
 *                           an independent long-double brute force over the 15
 *                           features where the reference's selection is sound,
 *                           closed forms (vertex over face), hand-worked plane
 *                           pre-pass codes, both case-1 fall-throughs and the
 *                           early returns on zero normals.
 *   or_adaptive_*           hand-worked head-on cases (approach, separation,
 *                           the close proposal and each branch of the update)
 *                           in tests/test_oracle_adaptive.py.
 *   or_ispc_pen             or_zoltan_penalty's setup, step and output
 *                           (golden-pinned) with its own schedule, and an
 *                           independent numpy Gauss-Newton formulation of the
 *                           5 regularised steps (dense 4x4 solve).
 */
#ifndef DELTA_ORACLE_H
#define DELTA_ORACLE_H

#include <stdint.h>
#include "../include/delta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bf.cpp:217-1178.  Returns 1 if the edge-edge result was kept, 0 otherwise. */
int or_bf_solver(const double* xA, const double* yA, const double* zA,
                 const double* xB, const double* yB, const double* zB,
                 double P[3], double Q[3]);

/* penalty.cpp:278-431 (four fixed Newton steps; `failed` from the last). */
void or_penalty_solver(const double* xA, const double* yA, const double* zA,
                       const double* xB, const double* yB, const double* zB,
                       double P[3], double Q[3], double max_err, int* failed);

/* Same solver from an arbitrary start x0[4] (the learned warm start). */
void or_penalty_solver_from(const double* xA, const double* yA, const double* zA,
                            const double* xB, const double* yB, const double* zB,
                            const double x0[4], int iterations,
                            double P[3], double Q[3], double max_err, int* failed,
                            double x_out[4]);

/* penalty.cpp:434-581 (adaptive, <= 16 steps).  *its is only written when the
 * iteration converges, as in the reference. */
void or_penalty_solver_adaptive(const double* xA, const double* yA, const double* zA,
                                const double* xB, const double* yB, const double* zB,
                                double P[3], double Q[3], double max_err, int* its);

/* contactpoint.cpp:116-155. */
void or_contact_make(double xPA, double yPA, double zPA, double epsA, int idA,
                     double xQB, double yQB, double zQB, double epsB, int idB,
                     int friction, dh_contact* c);

/* bf / penalty / hybrid* / *Stat drivers on two particles. */
int or_detect(int method,
              const double* xA, const double* yA, const double* zA, int nA,
              double epsA, int fricA, int idA,
              const double* xB, const double* yB, const double* zB, int nB,
              double epsB, int fricB, int idB,
              dh_contact* out, int cap, int* n_out, dh_stats* stats);

/* Unique-pair form (soa[k*n+i] layout of dh_tri_pairs).  method is one of
 * DH_PAIR_BF64, DH_PAIR_PENALTY64, DH_PAIR_HYBRID64, DH_PAIR_PENALTY_ADAPT64. */
int or_tri_pairs(int method, const double* soa, int64_t n, double eps_total,
                 double* P, double* Q, double* d, uint8_t* flags);

/* forces.cpp:291-396 for one particle pair, accumulating into force/torque. */
void or_contacts_forces(const dh_contact* c, int nc, int negate_normal,
                        const double posA[3], const double angA[3], const double linA[3],
                        double massA,
                        const double posB[3], const double angB[3], const double linB[3],
                        double massB,
                        const dh_force_params* fp, double force[3], double torque[3]);

/* dynamics.cpp:26-102, 104-138, 140-213, 215-242. */
void or_expmap(double o1, double o2, double o3, double L[9]);
void or_update_rotation_matrix(double angular[3], const double refAngular[3],
                               double rotation[9], double step);
void or_update_angular(double refAngular[3], const double rotation[9],
                       const double inertia[9], const double inverse[9],
                       const double torque[3], double step);
void or_update_vertex(double* x, double* y, double* z, double rx, double ry, double rz,
                      const double rotation[9], const double position[3],
                      const double refposition[3]);

/* One explicit DEM step over a scene (Engine.cpp:155-163 order: contact
 * detection on the current geometry, forces, velocity/angular update, gravity,
 * position/rotation/vertex update).  Candidate pairs are all (i<j) whose
 * AABBs inflated by eps overlap and that are not both obstacles.  Obstacles
 * with a velocity move kinematically (the engine's extension for rotating
 * boxes; pinned by a rigid-rotation closed form in the tests).
 * Returns the number of contacts. */
typedef struct or_scene {
  int n;                 /* particles                                 */
  int* tri_offset;       /* n+1                                       */
  double* x; double* y; double* z;          /* spatial vertices      */
  double* rx; double* ry; double* rz;       /* referential vertices  */
  double* eps; int* friction; int* id;
  dh_bodies b;           /* host pointers                             */
} or_scene;

int64_t or_step(or_scene* s, int method, const dh_force_params* fp, double dt,
                const double gravity[3], int64_t* checks, int* npairs_out);

/* zoltan/bf.cpp:60-952; -1 when the reference returns before writing P, Q. */
int or_zoltan_bf(const double* xA, const double* yA, const double* zA,
                 const double* xB, const double* yB, const double* zB,
                 double P[3], double Q[3]);
/* zoltan/penalty.cpp:27-208 and zoltan/hybrid.cpp:37-132 (serial semantics). */
void or_zoltan_penalty(unsigned it, unsigned nt, const double a[3], const double b[3],
                       const double c[3], double* const d[3], double* const e[3],
                       double* const f[3], double* const p[3], double* const q[3]);
void or_zoltan_hybrid(int it, int nt, const double a[3], const double b[3], const double c[3],
                      double* const d[3], double* const e[3], double* const f[3],
                      double* const p[3], double* const q[3]);
/* zoltan/contact.cpp:56-240 (see dh_zoltan_detection); all != 0: j > i in
 * [s1, e1), else j in [s2, e2).  Returns the number of kept contacts. */
int or_zoltan_detection(int s1, int e1, int s2, int e2, int all, double* const t[6][3],
                        const int* pid, dh_zoltan_contact* out, int cap);

/* ISPC kernels bf.ispc:3-937 and penalty.ispc:3-174 (REAL = double) over
 * vertex-major SoA tables pp[v][k][i]; P, Q, distance in/out as in ispc_bf. */
void or_ispc_bf(unsigned n, double* const pp[6][3], double* const P[3], double* const Q[3],
                double* distance);
void or_ispc_pen(unsigned n, double* const pp[6][3], double* const P[3], double* const Q[3],
                 double* distance);

/* Adaptive time step (multiscale/dem/mappings/Collision.cpp:188-299,
 * State.cpp:26-35,137-199): the pairs at the start of a step, then the
 * update after it (see dh_adaptive_state). */
void or_adaptive_pairs(const int* pairs, int npairs, const double* pos, const double* linvel,
                       const double* radius, const double* eps, dh_adaptive_state* s);
void or_adaptive_finish(const double* linvel, const int* obstacle, int n, dh_adaptive_state* s);

/* GJK distance between two convex point sets (gjk.cpp:302-460). */
double or_gjk(const double* a, int na, const double* b, int nb, double p[3], double q[3]);

#ifdef __cplusplus
}
#endif
#endif
