/*
 * delta_hip.h — C ABI of the MI355X contact-detection and time-stepping engine.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns a status code (DH_OK == 0).  Pointers documented as "device" must be
 * HBM allocations made with dh_malloc (or any hipMalloc'd memory of the same
 * process); "host" pointers are ordinary process memory.  No entry point takes
 * ownership of caller buffers.
 *
 * Geometry convention (same as the reference): a particle is a flat triangle
 * soup stored as three arrays x[], y[], z[] of length 3*ntri, where
 * x[3*t+k] is the x coordinate of vertex k of triangle t
 * (delta/geometry/mesh/Mesh.cpp:248-302, delta/core/data/ParticleRecord.h:78-84).
 * All geometry arithmetic of the exact paths is IEEE binary64 without fused
 * multiply-add, in the reference's operation order.
 */
#ifndef DELTA_HIP_H
#define DELTA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
  DH_OK = 0,
  DH_ERR_ARG = 1,        /* null pointer, negative size, unknown method      */
  DH_ERR_HIP = 2,        /* a HIP runtime call failed (see dh_last_error)    */
  DH_ERR_CAPACITY = 3,   /* more contacts than `cap`; *n_out holds the need  */
  DH_ERR_REF_UB = 4,     /* input on which the reference reads out of bounds */
  DH_ERR_NO_DEVICE = 5,  /* no GPU visible                                   */
  DH_ERR_UNSUPPORTED = 6 /* method not available for this entry point        */
};

/* ---------------------------------------------------------------- methods */
/* Particle-pair detectors: one value per reference free function. */
enum {
  DH_BF = 0,                     /* detection::bf              bf.cpp:28-149      */
  DH_PENALTY = 1,                /* detection::penalty         penalty.cpp:113-241 */
  DH_HYBRID_BATCH = 2,           /* hybridWithPerBatchFallBack hybrid.cpp:338-487  */
  DH_HYBRID_TRIANGLE = 3,        /* hybridWithPerTriangleFallBack hybrid.cpp:59-200 */
  DH_PENALTY_STAT = 4,           /* penaltyStat                penalty.cpp:50-111  */
  DH_HYBRID_BATCH_STAT = 5,      /* hybridBatchStat            hybrid.cpp:489-627  */
  DH_HYBRID_TRIANGLE_STAT = 6,   /* hybridTriangleStat         hybrid.cpp:202-336  */
  DH_GJK = 7,                    /* detection::gjk             gjk.cpp:25-100      */
  DH_BF_LEARNED = 8              /* DH_BF's contacts bit for bit; the survivor queue's
                                    triangle pairs (always on) go through the learned
                                    narrow phase (network of dh_learned_set_model,
                                    features in FP64, support test), the rest through
                                    bfSolver.  No counterpart in the reference.  */
};

/* Unique triangle-pair kernels (the isolated narrow-phase workload). */
enum {
  DH_PAIR_BF64 = 0,        /* bfSolver, FP64, bit-exact                            */
  DH_PAIR_PENALTY64 = 1,   /* penaltySolver (4 fixed Newton steps), FP64, bit-exact */
  DH_PAIR_HYBRID64 = 2,    /* penalty, bfSolver where the pair's `failed` is set   */
  DH_PAIR_MIXED = 3,       /* FP32 feature screen + FP64 exact finalisation        */
  DH_PAIR_LEARNED = 4,     /* MLP feature guess + FP64 feature evaluation, support
                              check and tied features folded; the rest bfSolver */
  DH_PAIR_PENALTY_ADAPT64 = 5, /* penaltySolver(..., int& its), <=16 its, FP64      */
  DH_PAIR_BF64_EXACT = 6,  /* segment-triangle intersection test first (the zoltan
                              segt of zoltan/hybrid.cpp:280-342, in its order of
                              hybrid.cpp:652-670): intersecting pairs get P = Q =
                              the crossing point, d = 0; the rest bfSolver      */
  DH_PAIR_LEARNED_PENALTY = 7, /* the network's barycentric outputs start the penalty
                                  Newton iteration (<= 16 steps, residual 1e-10), the
                                  result is clamped onto the triangles and certified
                                  like DH_PAIR_LEARNED; the rest bfSolver          */
  DH_PAIR_PENALTY_CERT = 8     /* the same from the reference's start x = 0.33
                                  (penalty.cpp:353-356), no network: what the
                                  network's start gains                           */
};

/* flags[] bits written by dh_tri_pairs */
enum {
  DH_FLAG_FAILED = 1,      /* penalty `failed` bit of the last Newton step         */
  DH_FLAG_RECHECK = 2,     /* FP64 path taken (screen could not certify)           */
  DH_FLAG_EDGE = 4,        /* edge-edge feature won (ptmin > ssmin)                */
  DH_FLAG_CERTIFIED = 8,   /* learned guess certified without the full BF          */
  DH_FLAG_INTERSECT = 16   /* DH_PAIR_BF64_EXACT: the triangles intersect, d = 0   */
};

/* ------------------------------------------------------------ contact record */
/* Mirror of delta::contact::contactpoint (contactpoint.h:42-103) plus the
 * triangle indices and the batched-pair index, so that contact sets compare
 * as sets of (master, slave, triA, triB). */
typedef struct dh_contact {
  double x[3];        /* (P+Q)/2                                           */
  double normal[3];   /* (P-Q)/|P-Q|                                        */
  double depth;       /* eps_total - |P-Q|                                  */
  double P[3];        /* closest point on the triangle of particle A        */
  double Q[3];        /* closest point on the triangle of particle B        */
  double eps_total;   /* epsA + epsB                                        */
  int32_t master;     /* id of particle A                                   */
  int32_t slave;      /* id of particle B                                   */
  int32_t penetration;/* always 1 for the 11-argument constructor           */
  int32_t friction;   /* frictionA && frictionB                             */
  int32_t triA;       /* triangle index within A                            */
  int32_t triB;       /* triangle index within B                            */
  int32_t pair;       /* index into the particle-pair list of the call      */
  int32_t pad_;
} dh_contact;

/* Statistics the reference keeps in static globals (penalty.cpp:33-48,
 * hybrid.cpp:27-57).  The caller owns and zeroes it; calls accumulate. */
typedef struct dh_stats {
  int64_t newton_histogram[17]; /* penaltyStat: _numberOfNewtonIterations     */
  int64_t penalty_fails;        /* _numberOfPenaltyFails                      */
  int64_t batch_fails;          /* _numberOfBatchFails                        */
  int64_t batch_size;           /* _batchSize                                 */
  double batch_error;           /* _batchError                                */
  int64_t triangle_checks;      /* nTriA*nTriB per particle pair (Engine.cpp:451) */
  int64_t pairs_evaluated;      /* triangle-pair evaluations of the reference */
  int64_t fallback_rows;        /* rows recomputed with bfSolver              */
  int64_t pairs_culled;         /* of those, DH_BF pairs not solved: their    */
                                /* boxes grown by the margin are apart, so no */
                                /* contact (no counterpart in the reference)  */
} dh_stats;

/* ------------------------------------------------------------ housekeeping */
int dh_init(void);                          /* selects device 0, idempotent  */
const char* dh_last_error(void);
int dh_device_count(int* n);
int dh_malloc(void** dev_ptr, size_t bytes);
int dh_free(void* dev_ptr);
int dh_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int dh_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int dh_memset(void* dev_ptr, int value, size_t bytes);
int dh_synchronize(void);
/* Free and total HBM of the device (hipMemGetInfo), for footprint reports. */
int dh_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Timing on the library stream (or a caller stream when `stream` != NULL). */
int dh_event_create(void** ev);
int dh_event_destroy(void* ev);
int dh_event_record(void* ev, void* stream);
int dh_event_elapsed_ms(float* ms, void* start, void* stop);

/* ------------------------------------------------ particle-pair detection */
/* Host-pointer form of delta::contact::detection::{bf, penalty, hybrid*,
 * penaltyStat, hybrid*Stat} (bf.h:65-81, penalty.h:62-115, hybrid.h:57-170).
 * nTriA/nTriB are triangle counts.  Contacts are written to `out` (host) in
 * the reference's emission order (row iA outer, iB inner); `*n_out` is the
 * number of contacts.  If more than `cap` would be written, DH_ERR_CAPACITY
 * is returned and *n_out holds the required capacity.  `stats` may be NULL. */
int dh_detect_pair(int method,
                   const double* xA, const double* yA, const double* zA,
                   int nTriA, double epsA, int frictionA, int idA,
                   const double* xB, const double* yB, const double* zB,
                   int nTriB, double epsB, int frictionB, int idB,
                   dh_contact* out, int cap, int* n_out, dh_stats* stats);

/* Device-resident batch over particle pairs.  Geometry of all particles is
 * concatenated: triangles of particle p are [tri_offset[p], tri_offset[p+1]).
 * pairs[2*k], pairs[2*k+1] are particle indices (A, B).  Contacts of pair k
 * are written contiguously in reference order, pairs in ascending k, to
 * out_dev; *n_out (host) receives the count.  pair_count_dev (may be NULL)
 * receives the per-pair contact count (device, npairs ints). */
typedef struct dh_geometry {
  const double* x;          /* device, 3*ntri_total */
  const double* y;
  const double* z;
  const int* tri_offset;    /* device, nparticles+1 */
  const double* eps;        /* device, nparticles   */
  const int* friction;      /* device, nparticles   */
  const int* id;            /* device, nparticles   */
  int nparticles;
  int ntri_total;
} dh_geometry;

int dh_detect_pairs(int method, const dh_geometry* geom,
                    const int* pairs_dev, int npairs,
                    dh_contact* out_dev, int cap, int* n_out,
                    int* pair_count_dev, dh_stats* stats, void* stream);

/* -------------------------------------------------- unique triangle pairs */
/* soa_dev: 18 device arrays of n doubles, packed as soa_dev[k*n + i] with
 * k = 0..2 x of A's vertices 0..2, 3..5 y of A, 6..8 z of A, 9..17 the same
 * for B (the bfSolver argument layout, bf.h:95-108).  Outputs (device):
 * P[3*n] and Q[3*n] (component-major: P[c*n+i]), d[n] = |Q-P| as bf.cpp:117,
 * flags[n] (may be NULL).  eps_total is the contact threshold epsA+epsB used by
 * the penalty tolerance (eps_total/16, penalty.cpp:198) and by the mixed path's
 * recheck band.  soa32_dev (may be NULL) is the same data rounded to binary32,
 * read by DH_PAIR_MIXED / DH_PAIR_LEARNED for the screening pass. */
int dh_tri_pairs(int method, const double* soa_dev, const float* soa32_dev,
                 int64_t n, double eps_total,
                 double* P_dev, double* Q_dev, double* d_dev, uint8_t* flags_dev,
                 void* stream);

/* Seeded synthetic unique-pair workload (no counterpart in the reference; the
 * isolated narrow-phase benchmark input).  A: vertices uniform in [0,box)^3.
 * Separation s log-uniform in [sep_min, sep_max].  frac_parallel of the pairs
 * put an edge of B at angle < max_angle to an edge of A, at distance s, with
 * those edges as the closest features; frac_intersect are pushed through A;
 * the rest translate a random B along a random axis until the true distance
 * is s (bisection with bisection_steps steps).  Writes soa_dev (18*n, the
 * dh_tri_pairs layout), target_dev[n] = s and kind_dev[n] in {0 general,
 * 1 near-parallel, 2 intersecting} (both may be NULL).  p may be NULL for
 * the defaults (1e-3, 1e-7, 1e-2, 0.05, 0.01, 1e-4, 48). */
typedef struct dh_gen_params {
  double box, sep_min, sep_max, frac_parallel, frac_intersect, max_angle;
  int bisection_steps;
} dh_gen_params;
int dh_gen_tri_pairs(uint64_t seed, int64_t n, const dh_gen_params* p, double* soa_dev,
                     double* target_dev, uint8_t* kind_dev, void* stream);

/* Element-wise binary64 -> binary32 rounding of n values (device). */
int dh_convert_f32(const double* src_dev, float* dst_dev, int64_t n, void* stream);

/* ------------------------------------------- zoltan application kernels */
/* zoltan/bf.cpp:60-952 (`bf(A3x, A3y, A3z, B3x, B3y, B3z, &P, &Q)`) for n unique
 * pairs in the dh_tri_pairs layout.  bfSolver's arithmetic with the zoltan
 * differences: a pair whose normal test returns early (zoltan/bf.cpp:111-119,
 * 165-169) keeps its P/Q/d untouched and gets DH_FLAG_FAILED; the switch
 * fall-throughs of cases 1 are kept.  d (may be NULL) = |Q-P|. */
int dh_zoltan_bf(const double* soa_dev, int64_t n, double* P_dev, double* Q_dev,
                 double* d_dev, uint8_t* flags_dev, void* stream);

/* zoltan/penalty.cpp:27-208 (`penalty(it, nt, a, b, c, d, e, f, p, q)`):
 * triangle (a, b, c) (host) against triangles i in [it, nt) of the
 * vertex-major device arrays d[k][i], e[k][i], f[k][i] (host tables of three
 * device pointers); writes p[k][i], q[k][i].  Five Newton steps; from i = 4 on
 * the regularisation is multiplied by r before every step, as in the
 * reference (penalty.cpp:153, which tests the pair index). */
int dh_zoltan_penalty(unsigned it, unsigned nt, const double a[3], const double b[3],
                      const double c[3], double* const d[3], double* const e[3],
                      double* const f[3], double* const p[3], double* const q[3],
                      void* stream);

/* zoltan/hybrid.cpp:37-132 (`hybrid(it, nt, ...)`), same arrays: batches k
 * in [it, nt/4) of the four pairs 4k..4k+3; seven-step penalty per pair; when
 * the batch's mean error exceeds 1e-8, the brute force of
 * zoltan/hybrid.cpp:646-885 on the batch's last pair is written to all four
 * (the reference's serial semantics: its fallback reuses the last pair's
 * D, E, F). */
int dh_zoltan_hybrid(int it, int nt, const double a[3], const double b[3], const double c[3],
                     double* const d[3], double* const e[3], double* const f[3],
                     double* const p[3], double* const q[3], void* stream);

/* zoltan/contact.h:34-43 contact record. */
typedef struct dh_zoltan_contact {
  int pid[2];
  int color[2];
  double point[3];
  double normal[3];   /* (Q - P) / depth: not unit length (contact.cpp:120-122) */
  double depth;       /* 0.15 - |Q - P|                                      */
  double pp[3];
  double qq[3];
} dh_zoltan_contact;

/* zoltan/contact.cpp:56-146 (`contact::detection(s, e, t, tid, pid, ...)`):
 * every triangle pair i < j in [s, e) of different particles, t[v][axis][i]
 * the spatial vertices (host table of device pointers; t[3..5] unused), pid
 * (device) in [0, npid).  A pair is a contact when zoltan's bf gives
 * |Q - P| < 0.15; it is dropped when any coordinate of its P or Q equals that
 * of a contact its particle pid[i] already holds from this call.  out (device)
 * receives the kept contacts in discovery order (i, then j); the reference's
 * per-particle list conpnt[p] is the subsequence with pid[0] == p.  Differs
 * from the reference where that is undefined: triangle i's y and z are read
 * (the reference reads them into By, Bz and leaves Ay, Az uninitialised,
 * contact.cpp:66-72), and pairs for which bf returns early are skipped. */
int dh_zoltan_detection(int s, int e, double* const t[6][3], const int* tid_dev,
                        const int* pid_dev, int npid, dh_zoltan_contact* out_dev, int cap,
                        int* n_out, void* stream);

/* zoltan/contact.cpp:149-240: the same for i in [s1, e1) against j in [s2, e2). */
int dh_zoltan_detection2(int s1, int e1, int s2, int e2, double* const t[6][3],
                         const int* tid_dev, const int* pid_dev, int npid,
                         dh_zoltan_contact* out_dev, int cap, int* n_out, void* stream);

/* ------------------------------------------------------- ISPC kernels */
/* ispc_bf (delta/contact/detection/bf.ispc:3-6), REAL = double: n pairs,
 * ppV[k] = device array of coordinate k of vertex V (pp1..pp3 triangle A,
 * pp4..pp6 triangle B).  P, Q and distance are in/out, as in the reference:
 * - a zero normal of B, or the mixed normal test of bf.ispc:110, writes
 *   distance = -1 and leaves P and Q;
 * - otherwise a plane pre-pass writes a code (2 segment in plane, -2 no
 *   crossing, -1 crossing outside, 0 crossing inside) and its crossing point;
 *   the nine edge-edge solves (SMALL_NUM 1e-7) overwrite P and Q, and when
 *   the smallest vertex-face distance is below the edge-edge one it replaces
 *   distance, P and Q; otherwise distance keeps the code;
 * - both output switches fall through from case 1 into case 2 (edge pair 1
 *   takes its Q from B's edge 2-0, vertex-face 1 reports B's vertex 2). */
int dh_ispc_bf(unsigned n, double* const pp1[3], double* const pp2[3], double* const pp3[3],
               double* const pp4[3], double* const pp5[3], double* const pp6[3],
               double* const P[3], double* const Q[3], double* distance, void* stream);

/* ispc_pen (penalty.ispc:154-174): five Newton steps of the zoltan penalty
 * (eps 0.1, r = lambda 10^(ln lambda + 14)), the regularisation multiplied by
 * r before the fifth step; writes p, q and distance = |p - q|. */
int dh_ispc_pen(unsigned n, double* const a[3], double* const b[3], double* const c[3],
                double* const d[3], double* const e[3], double* const f[3],
                double* const p[3], double* const q[3], double* distance, void* stream);

/* ------------------------------------------------- learned narrow phase */
/* Feature network (no counterpart in the reference): 18 canonical-frame
 * coordinates -> 64 -> 64 -> 32 outputs (15 closest-feature logits, an
 * exact-path logit, 4 barycentric parameters), ReLU, bf16 MFMA with FP32
 * accumulation.  Parameters are dh_mlp_param_count() floats (device).
 * dh_mlp_labels: for n pairs (dh_tri_pairs layout) writes the network input
 * X (n x 32 bf16), bfSolver's winning feature (0..14), the barycentric
 * parameters of its P and Q (4n floats) and the exact-path label (1 when
 * bfSolver's own result cannot be certified: intersecting or degenerate).
 * dh_mlp_grad: loss and gradient over pairs [offset, offset+batch) (grad
 * overwritten).  dh_mlp_adam: one Adam step (step >= 1).
 * dh_learned_set_model: copies the parameters DH_PAIR_LEARNED uses. */
int dh_mlp_param_count(void);
int dh_mlp_labels(const double* soa_dev, int64_t n, uint16_t* X_dev, uint8_t* feat_dev,
                  float* bary_dev, float* flag_dev, void* stream);
int dh_mlp_forward(const float* params_dev, const uint16_t* X_dev, int64_t n, float* out_dev,
                   void* stream);
int dh_mlp_grad(const float* params_dev, const uint16_t* X_dev, const uint8_t* feat_dev,
                const float* bary_dev, const float* flag_dev, int64_t offset, int batch,
                float* grad_dev, float* loss_dev, void* stream);
int dh_mlp_adam(float* params_dev, float* m_dev, float* v_dev, const float* grad_dev, float lr,
                int step, void* stream);
int dh_learned_set_model(const float* params_dev);

/* Number of pairs of the last dh_tri_pairs call per path (host). */
typedef struct dh_pair_counters {
  int64_t screened;     /* pairs finished by the FP32 screen (MIXED)        */
  int64_t rechecked;    /* pairs finished by the full FP64 bfSolver         */
  int64_t features_fp64;/* FP64 feature evaluations (LEARNED) or Newton steps
                           (LEARNED_PENALTY, PENALTY_CERT)                    */
  int64_t certified;    /* pairs certified from the learned guess           */
} dh_pair_counters;
int dh_tri_pairs_counters(dh_pair_counters* out);

/* ------------------------------------------------------------- autotuning */
/* No counterpart in the reference (its kernels have fixed ISPC task and
 * OpenMP schedules).  Launch parameters of the hot kernels, process-wide:
 * they change how the work is spread over the GPU, never the results (every
 * setting gives bit-identical outputs).  Read at each launch; a captured
 * step graph keeps the values of its creation. */
typedef struct dh_tuning {
  int32_t learned_wg_per_cu;  /* learned passes: persistent workgroups per CU
                                 (1..32); also sizes their failure queues   */
  int32_t mixed_wg_per_cu;    /* DH_PAIR_MIXED: the same (1..32)             */
  int32_t exact_block;        /* FP64 one-pair-per-lane kernels of
                                 dh_tri_pairs: 64, 128, 192 or 256           */
  int32_t detect_tile_target; /* particle-pair detection: workgroups per pair
                                 = clamp(ceil(target / npairs), 1, max)      */
  int32_t detect_max_tiles;   /* that max (1..64)                            */
  int32_t detect_wg_per_cu;   /* particle-pair detection: 0 = one workgroup
                                 per (pair, rows) item; else a persistent grid
                                 of this many workgroups per CU (1..64)      */
  int32_t detect_queue_min_items; /* DH_BF detection of at least this many
                                 items (pairs x tiles) solves the survivors of
                                 the box test from one queue, one per lane
                                 (0..2^30)                                   */
  int32_t detect_fp32_screen; /* 1: queued survivors are first solved in
                                 binary32 (coordinates relative to A's first
                                 vertex); those at twice the margin or more
                                 are dropped, the rest solved in binary64, so
                                 every contact is still decided and computed
                                 in binary64 (0 or 1)                        */
} dh_tuning;
int dh_tuning_default(dh_tuning* out);
int dh_tuning_get(dh_tuning* out);
int dh_tuning_set(const dh_tuning* t);  /* DH_ERR_ARG if a field is out of range */

/* One timed candidate: knob, value, median milliseconds per call. */
enum {
  DH_KNOB_LEARNED_WG_PER_CU = 0,
  DH_KNOB_MIXED_WG_PER_CU = 1,
  DH_KNOB_EXACT_BLOCK = 2,
  DH_KNOB_DETECT_TILE_TARGET = 3,
  DH_KNOB_DETECT_MAX_TILES = 4,
  DH_KNOB_DETECT_WG_PER_CU = 5,
  DH_KNOB_DETECT_QUEUE_MIN_ITEMS = 6,
  DH_KNOB_DETECT_FP32_SCREEN = 7
};
typedef struct dh_tune_trial {
  int32_t knob;
  int32_t value;
  float ms;
} dh_tune_trial;

/* Times dh_tri_pairs(method, ...) on the caller's pairs for every candidate
 * of the knob that governs `method` (exact methods: exact_block; MIXED:
 * mixed_wg_per_cu; learned methods: learned_wg_per_cu): one untimed round
 * over the candidates, then `reps` timed rounds in rotated order (HIP events
 * on `stream`, one call each), and keeps the candidate of least median time
 * in the table.  trials[0..trial_cap) receive the candidates, *ntrials their
 * number.  On return the outputs hold the chosen configuration's results. */
int dh_autotune_tri_pairs(int method, const double* soa_dev, const float* soa32_dev, int64_t n,
                          double eps_total, double* P_dev, double* Q_dev, double* d_dev,
                          uint8_t* flags_dev, int reps, dh_tune_trial* trials, int trial_cap,
                          int* ntrials, void* stream);
/* The same for dh_detect_pairs: sweeps detect_tile_target (with the cap at
 * 64), then detect_max_tiles, then detect_wg_per_cu, then the survivor
 * queue on (0) or off (2^30) as detect_queue_min_items, then its binary32
 * screen (detect_fp32_screen 0, 1), timing the detection
 * as dh_step runs it (with the particle boxes of the broad phase).  `cap`
 * must hold every contact. */
int dh_autotune_detect(int method, const dh_geometry* geom, const int* pairs_dev, int npairs,
                       dh_contact* out_dev, int cap, int reps, dh_tune_trial* trials,
                       int trial_cap, int* ntrials, void* stream);

/* ------------------------------------------------------- forces / dynamics */
/* Rigid-body state, one record per particle (device, AoS of small arrays). */
typedef struct dh_bodies {
  double* pos;        /* 3n  spatial centre of mass      (_centreOfMass)      */
  double* refpos;     /* 3n  referential centre of mass  (_refCentreOfMass)   */
  double* linvel;     /* 3n  linear velocity                                  */
  double* angvel;     /* 3n  spatial angular velocity    (_angularVelocity)   */
  double* refangvel;  /* 3n  referential angular velocity                     */
  double* rot;        /* 9n  orientation, column-major (NVMUL, algo.h:980)    */
  double* inertia;    /* 9n  referential inertia                              */
  double* inverse;    /* 9n  its inverse                                      */
  double* mass;       /* n   (may be +inf for obstacles)                      */
  int* obstacle;      /* n   1: no force moves it; at rest it stays as built,
                              with a velocity it moves kinematically          */
  int* material;      /* n                                                    */
  int n;
} dh_bodies;

/* delta::contact::forces::getContactsForces (forces.cpp:291-396) for a batch
 * of particle pairs whose contacts were produced by dh_detect_pairs.
 * For pair k, contacts [pair_offset[k], pair_offset[k+1]) of contacts_dev are
 * applied to A = pairs[2k] (normal as stored) and to B = pairs[2k+1] (normal
 * negated, Collision.cpp:177-185).  Per-particle force/torque (device, 3n
 * each) are overwritten with the sum over the particle's pairs, in ascending
 * partner index, of the per-pair sums in contact order.  touched_dev (n ints,
 * may be NULL) receives 1 for particles with at least one contact: only those
 * get the force/angular update (Engine.cpp:464).  Constants: spring k,
 * damping ratio, friction coefficient (is_sphere: the caller passes the
 * springSphere constants SSPRING, SDAMPER, SFRICTIONGOLD, forces.cpp:17-19). */
typedef struct dh_force_params {
  double spring;      /* SPRING 1E4   (forces.cpp:11) */
  double damper;      /* DAMPER 0.5   (forces.cpp:12) */
  double friction;    /* FRICTION 0.5 (forces.cpp:14) */
  int is_sphere;
} dh_force_params;

int dh_contacts_forces(const dh_bodies* bodies, const int* pairs_dev, int npairs,
                       const dh_contact* contacts_dev, const int64_t* pair_offset_dev,
                       const dh_force_params* fp,
                       double* force_dev, double* torque_dev, int* touched_dev, void* stream);

/* Engine::deriveForces tail + Engine::updatePosition (Engine.cpp:511-565):
 * for touched particles v += dt F/m and updateAngular; then for every
 * non-obstacle particle v += dt*g (non-zero components), x += dt v,
 * updateRotationMatrix, and every vertex replica x = c + R (X - C_ref).
 * An obstacle with a non-zero linear or referential angular velocity (no
 * counterpart in the reference, whose obstacles never move) takes the same
 * x, R and vertex update without force or gravity: a rotating box or drum.
 * touched_dev may be NULL (all particles touched). */
int dh_integrate(dh_bodies* bodies, const double* force_dev, const double* torque_dev,
                 const int* touched_dev, double dt, const double gravity[3],
                 double* x_dev, double* y_dev, double* z_dev,
                 const double* refx_dev, const double* refy_dev, const double* refz_dev,
                 const int* tri_offset_dev, void* stream);

/* Host-pointer single-particle forms of the dynamics free functions
 * (dynamics.h:45-95), for drop-in callers. */
int dh_update_angular(double refAngular[3], const double rotation[9],
                      const double inertia[9], const double inverse[9],
                      const double torque[3], double step);
int dh_update_rotation_matrix(double angular[3], const double refAngular[3],
                              double rotation[9], double step);

/* ------------------------------------------------------------ broad phase */
/* Candidate particle pairs (no counterpart function in the reference, whose
 * engine loops over all pairs, Engine.cpp:272-295): every pair (i<j), not both
 * obstacles, whose AABBs inflated by their margins (+0.5e-9 m each) overlap,
 * found with a uniform hash grid of cell size `cell` (scenes of at most 1024
 * particles: every pair tested in one workgroup, same result).  Obstacles
 * (obstacle_dev, n ints, may be NULL) and particles whose inflated box is
 * wider than `cell` are tested against all others; choose `cell` >= the
 * typical particle's inflated extent.  Writes pairs_dev[2k], pairs_dev[2k+1],
 * sorted ascending; *npairs receives the count (DH_ERR_CAPACITY if more than
 * cap; the first cap pairs may then have been written). */
int dh_broadphase(const dh_geometry* geom, const int* obstacle_dev, double cell,
                  int* pairs_dev, int cap, int* npairs, void* stream);

/* dh_broadphase without any allocation: no counterpart in the reference.
 * For callers that cannot call hipMalloc (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Same arguments,
 * results and errors as dh_broadphase; its working buffers live in one fixed
 * device array of the library (8,523,264 bytes per device, held until the
 * process ends) instead of library buffers that grow on demand.  Accepts at
 * most DH_NOALLOC_BP_PARTICLES particles and DH_NOALLOC_BP_TRIANGLES
 * triangles (DH_ERR_CAPACITY beyond).  Calls on one device share that array:
 * calls that overlap on one device race, and must be ordered by the caller
 * (one stream, or events).  The pair count is still read back to the host
 * (one stream synchronisation), as in dh_broadphase. */
#define DH_NOALLOC_BP_PARTICLES 16384
#define DH_NOALLOC_BP_TRIANGLES (1 << 20)
int dh_broadphase_noalloc(const dh_geometry* geom, const int* obstacle_dev, double cell,
                          int* pairs_dev, int cap, int* npairs, void* stream);
/* Bytes of that array a broad phase of (nparticles, ntri) uses (host only). */
size_t dh_broadphase_noalloc_scratch_bytes(int nparticles, int ntri);

/* Exclusive prefix sum of n ints into n+1 int64 offsets (offs[n] = total). */
int dh_exclusive_scan(const int* counts_dev, int n, int64_t* offs_dev, void* stream);

/* ------------------------------------------------------------- time step */
/* One explicit DEM step in the order of Engine::iterate (Engine.cpp:155-163,
 * 270-567): broad phase, narrow phase with `method` on the current geometry,
 * getContactsForces per pair, v += dt F/m and updateAngular for particles
 * with contacts, then gravity, position, rotation and vertex replicas for
 * every non-obstacle particle.  Buffers of the scene are device memory owned
 * by the caller; intermediate buffers are owned by the library and grow on
 * demand.  info (host, may be NULL) receives the step's counters. */
typedef struct dh_scene {
  dh_geometry geom;         /* geom.x/y/z: spatial vertices, updated in place */
  double* x;
  double* y;
  double* z;
  const double* refx;       /* referential vertices (ParticleRecord._ref*)    */
  const double* refy;
  const double* refz;
  dh_bodies bodies;
} dh_scene;

/* Adaptive time-step state of the Peano runner (multiscale/dem/State.cpp,
 * mappings/Collision.cpp:188-299), kept in device memory so that captured
 * steps follow it.  The caller fills dt, min_dt, increase, mesh_width,
 * max_dt (+inf to start), max_vel_travel (0) and step (0); the rest is
 * written by the step.  Each step integrates with `dt` and leaves the next
 * step's `dt` there. */
typedef struct dh_adaptive_state {
  double dt;               /* step size of the next step                      */
  double time;             /* simulated time                                  */
  double min_dt;           /* mindt, State.cpp:156 (1e-4 there)               */
  double increase;         /* increaseFactor and StepIncrement (1.1)          */
  double mesh_width;       /* getMinMeshWidth: the broad-phase cell here      */
  double max_dt;           /* _maxdt of the last update                       */
  double close;            /* TwoParticlesAreClose of the last step, 0: none  */
  double max_vel_approach; /* MaxVelocityApproach of the last step            */
  double max_vel_travel;   /* fastest particle after the last step            */
  int64_t step;            /* steps taken                                     */
  int32_t separate;        /* some pair separated in the last step            */
  int32_t reserved_;
  uint64_t scratch[4];     /* reductions of the running step (library use)   */
} dh_adaptive_state;

typedef struct dh_step_params {
  int method;               /* DH_BF, DH_PENALTY, DH_HYBRID_*                 */
  double dt;                /* ignored when `adaptive` is set                 */
  double gravity[3];        /* reference: (0, -9.8, 0), Engine.cpp:534        */
  dh_force_params forces;
  double cell;              /* broad-phase cell size                          */
  dh_adaptive_state* adaptive; /* device; NULL: fixed dt (no counterpart on   */
                               /* the delta engine path, Peano's controller)  */
  const double* radius;     /* device, n: bounding radius of each particle    */
                            /* about its centre (needed with `adaptive`)      */
} dh_step_params;

/* The adaptive controller on its own (the step calls the same kernels):
 * phase 0 at the start of a step, on its candidate pairs (clears the
 * reductions, then triggerParticleTooClose for every pair); phase 1 after
 * the step's update (fastest particle, then finishedTimeStep /
 * adaptiveTimeStep / clearAccumulatedData).  No counterpart on the delta
 * engine path; mirrors multiscale/dem/State.cpp:122-199. */
int dh_adaptive_step(const int* pairs_dev, int npairs, const dh_bodies* bodies,
                     const double* radius_dev, const double* eps_dev,
                     dh_adaptive_state* state_dev, int phase, void* stream);

typedef struct dh_step_info {
  int64_t candidate_pairs;  /* broad-phase pairs                              */
  int64_t contacts;         /* contact points emitted                         */
  int64_t triangle_checks;  /* sum of nTriA*nTriB over candidates             */
  int64_t pairs_culled;     /* of those, skipped by the DH_BF box test        */
} dh_step_info;

int dh_step(dh_scene* scene, const dh_step_params* params, dh_step_info* info, void* stream);

/* The same step captured once into a HIP graph and replayed (no counterpart
 * in the reference: Engine::iterate, Engine.cpp:155-163, as one launch).
 * Capacities are fixed at creation: kernels are launched for pair_cap pairs
 * and contact_cap contacts and read the live counts from device memory, so a
 * replay needs no host synchronisation.  Methods DH_BF and DH_PENALTY.
 * dh_step_graph_status waits for the last replay (launch and status on the
 * same stream) and reports its counts; *overflow != 0 (bit 0 pairs >
 * pair_cap, bit 1 contacts > contact_cap, bit 2 a detection error) means
 * that some replay since the previous status call dropped work: the caller
 * restores its own copy of the state and recreates the graph with larger
 * capacities.
 * A graph keeps the scene pointers it was created with, and shares the
 * library's scratch with the other entry points.  dh_step_graph_launch
 * replays on `stream`, or on the graph's own blocking stream when NULL; a
 * non-blocking stream is rejected (DH_ERR_ARG), since it would not wait for
 * eager work on the null stream.  Each launch waits for the previous launch
 * of any graph, so replays of different graphs never overlap; eager entry
 * points called on a stream other than the null stream must be ordered with
 * replays by the caller. */
typedef struct dh_step_graph dh_step_graph;
int dh_step_graph_create(const dh_scene* scene, const dh_step_params* params, int64_t pair_cap,
                         int64_t contact_cap, dh_step_graph** out);
int dh_step_graph_launch(dh_step_graph* graph, void* stream);
int dh_step_graph_status(dh_step_graph* graph, void* stream, dh_step_info* info, int* overflow);
int dh_step_graph_destroy(dh_step_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* DELTA_HIP_H */
