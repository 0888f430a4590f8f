// tuning.hip — launch parameters of the hot kernels and the autotuner that
// chooses them on the device at hand.
//
// The table holds the knobs that change how work is laid over the GPU but
// never what is computed: the persistent grids of the learned and mixed
// passes (workgroups per compute unit, which also sizes their per-workgroup
// failure queues), the block size of the one-pair-per-lane FP64 kernels and
// the row tiling of the particle-pair detection (workgroups per pair).  Every
// candidate gives bit-identical outputs; the autotuner times each on the
// caller's own inputs with HIP events on the caller's stream and keeps the
// fastest in the table, where later launches read it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "runtime.hpp"

namespace dh {

static const dh_tuning kDefaults = {6, 8, 256, 4096, 16, 0, 32768, 0};
static dh_tuning g_tuning = kDefaults;
static int g_cus = 0;

const dh_tuning& tuning() { return g_tuning; }

int device_cus() {
  if (g_cus > 0) return g_cus;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    g_cus = cus;
  else
    g_cus = 256;  // MI355X
  return g_cus;
}

static int check_tuning(const dh_tuning& t) {
  if (t.learned_wg_per_cu < 1 || t.learned_wg_per_cu > 32)
    return set_error(DH_ERR_ARG, "dh_tuning: learned_wg_per_cu %d outside 1..32",
                     t.learned_wg_per_cu);
  if (t.mixed_wg_per_cu < 1 || t.mixed_wg_per_cu > 32)
    return set_error(DH_ERR_ARG, "dh_tuning: mixed_wg_per_cu %d outside 1..32",
                     t.mixed_wg_per_cu);
  if (t.exact_block < 64 || t.exact_block > 256 || t.exact_block % 64)
    return set_error(DH_ERR_ARG, "dh_tuning: exact_block %d is not 64, 128, 192 or 256",
                     t.exact_block);
  if (t.detect_tile_target < 1 || t.detect_tile_target > (1 << 20))
    return set_error(DH_ERR_ARG, "dh_tuning: detect_tile_target %d outside 1..2^20",
                     t.detect_tile_target);
  if (t.detect_max_tiles < 1 || t.detect_max_tiles > 64)
    return set_error(DH_ERR_ARG, "dh_tuning: detect_max_tiles %d outside 1..64",
                     t.detect_max_tiles);
  if (t.detect_queue_min_items < 0 || t.detect_queue_min_items > (1 << 30))
    return set_error(DH_ERR_ARG, "dh_tuning: detect_queue_min_items %d outside 0..2^30",
                     t.detect_queue_min_items);
  if (t.detect_fp32_screen != 0 && t.detect_fp32_screen != 1)
    return set_error(DH_ERR_ARG, "dh_tuning: detect_fp32_screen %d is not 0 or 1",
                     t.detect_fp32_screen);
  if (t.detect_wg_per_cu < 0 || t.detect_wg_per_cu > 64)
    return set_error(DH_ERR_ARG, "dh_tuning: detect_wg_per_cu %d outside 0..64",
                     t.detect_wg_per_cu);
  return DH_OK;
}

// Milliseconds of one call of `run` between two events on `st`.
template <typename F>
static int time_one(F& run, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float* ms) {
  DH_HIP_CHECK(hipEventRecord(e0, st));
  const int rc = run();
  if (rc) return rc;
  DH_HIP_CHECK(hipEventRecord(e1, st));
  DH_HIP_CHECK(hipEventSynchronize(e1));
  DH_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
  return DH_OK;
}

// Sweeps `values` of the knob `field` of the table and leaves the fastest
// in it (the table is restored on failure).  One untimed round over all
// candidates, then `reps` timed rounds, each starting one candidate further
// on, so that a drift of the clocks over the sweep (the GPU ramps up under
// sustained load) is spread over every candidate instead of favouring the
// last ones; a candidate's time is the median of its rounds.
template <typename F>
static int sweep(int knob, int dh_tuning::*field, const std::vector<int>& values, F run, int reps,
                 hipStream_t st, dh_tune_trial* trials, int trial_cap, int* ntrials) {
  const dh_tuning saved = g_tuning;
  const int nv = static_cast<int>(values.size());
  std::vector<std::vector<float>> t(nv);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DH_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = set_error(DH_ERR_HIP, "autotune: hipEventCreate failed");
  for (int k = 0; k < nv && rc == DH_OK; ++k) {
    g_tuning.*field = values[k];
    rc = run();
  }
  for (int r = 0; r < reps && rc == DH_OK; ++r)
    for (int c = 0; c < nv && rc == DH_OK; ++c) {
      const int k = (c + r) % nv;
      g_tuning.*field = values[k];
      float ms = 0.f;
      rc = time_one(run, st, e0, e1, &ms);
      t[k].push_back(ms);
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) {
    g_tuning = saved;
    return rc;
  }
  int best = -1;
  float best_ms = 0.f;
  for (int k = 0; k < nv; ++k) {
    std::sort(t[k].begin(), t[k].end());
    const float ms = t[k][t[k].size() / 2];
    if (trials && k < trial_cap) trials[k] = dh_tune_trial{knob, values[k], ms};
    if (best < 0 || ms < best_ms) {
      best = values[k];
      best_ms = ms;
    }
  }
  g_tuning = saved;
  g_tuning.*field = best;
  if (ntrials) *ntrials = nv;
  return DH_OK;
}

int detect_pairs_impl(int method, const dh_geometry* geom, const int* pairs, int npairs,
                      dh_contact* out, long long cap, long long* n_total, int* pair_count_user,
                      long long* offs_out, dh_stats* stats, hipStream_t st, const double* pbox,
                      const double* bbox, const double* psph);
int particle_boxes(const dh_geometry* geom, const double** out, hipStream_t st);
const double* broadphase_block_boxes();
const double* broadphase_spheres();

static bool is_exact(int method) {
  return method == DH_PAIR_BF64 || method == DH_PAIR_PENALTY64 || method == DH_PAIR_HYBRID64 ||
         method == DH_PAIR_PENALTY_ADAPT64 || method == DH_PAIR_BF64_EXACT;
}

static bool is_learned(int method) {
  return method == DH_PAIR_LEARNED || method == DH_PAIR_LEARNED_PENALTY ||
         method == DH_PAIR_PENALTY_CERT;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_tuning_default(dh_tuning* out) {
  if (!out) return set_error(DH_ERR_ARG, "dh_tuning_default: null");
  *out = kDefaults;
  return DH_OK;
}

int dh_tuning_get(dh_tuning* out) {
  if (!out) return set_error(DH_ERR_ARG, "dh_tuning_get: null");
  *out = g_tuning;
  return DH_OK;
}

int dh_tuning_set(const dh_tuning* t) {
  if (!t) return set_error(DH_ERR_ARG, "dh_tuning_set: null");
  const dh_tuning v = *t;
  const int rc = check_tuning(v);
  if (rc) return rc;
  g_tuning = v;
  return DH_OK;
}

int dh_autotune_tri_pairs(int method, const double* soa, const float* soa32, int64_t n,
                          double eps_total, double* P, double* Q, double* d, uint8_t* flags,
                          int reps, dh_tune_trial* trials, int trial_cap, int* ntrials,
                          void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (ntrials) *ntrials = 0;
  if (n <= 0 || !soa || !P || !Q || !d || reps < 1 || trial_cap < 0)
    return set_error(DH_ERR_ARG, "dh_autotune_tri_pairs: bad arguments");
  const hipStream_t st = as_stream(stream);
  auto run = [&] { return dh_tri_pairs(method, soa, soa32, n, eps_total, P, Q, d, flags, stream); };
  const std::vector<int> per_cu = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};
  if (is_exact(method))
    rc = sweep(DH_KNOB_EXACT_BLOCK, &dh_tuning::exact_block, {64, 128, 192, 256}, run, reps, st,
               trials, trial_cap, ntrials);
  else if (method == DH_PAIR_MIXED)
    rc = sweep(DH_KNOB_MIXED_WG_PER_CU, &dh_tuning::mixed_wg_per_cu, per_cu, run, reps, st,
               trials, trial_cap, ntrials);
  else if (is_learned(method))
    rc = sweep(DH_KNOB_LEARNED_WG_PER_CU, &dh_tuning::learned_wg_per_cu, per_cu, run, reps, st,
               trials, trial_cap, ntrials);
  else
    return set_error(DH_ERR_ARG, "dh_autotune_tri_pairs: unknown method %d", method);
  if (rc) return rc;
  // the outputs are those of the chosen configuration
  return run();
}

int dh_autotune_detect(int method, const dh_geometry* geom, const int* pairs_dev, int npairs,
                       dh_contact* out_dev, int cap, int reps, dh_tune_trial* trials,
                       int trial_cap, int* ntrials, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (ntrials) *ntrials = 0;
  if (!geom || npairs <= 0 || !pairs_dev || reps < 1 || trial_cap < 0)
    return set_error(DH_ERR_ARG, "dh_autotune_detect: bad arguments");
  const hipStream_t st = as_stream(stream);
  // timed as in a step: with the particle boxes of the broad phase, which
  // let the detection skip rows apart from the partner particle
  const double* pbox = nullptr;
  if ((rc = particle_boxes(geom, &pbox, st))) return rc;
  long long n_out = 0;
  auto run = [&] {
    const int r = detect_pairs_impl(method, geom, pairs_dev, npairs, out_dev, cap, &n_out,
                                    nullptr, nullptr, nullptr, st, pbox, broadphase_block_boxes(),
                                    broadphase_spheres());
    return r == DH_ERR_CAPACITY
               ? set_error(DH_ERR_CAPACITY, "dh_autotune_detect: %lld contacts, cap %d", n_out, cap)
               : r;
  };
  // tiles per pair = clamp(ceil(target / npairs), 1, max): the target is
  // swept with the cap at 64, then the cap with the chosen target
  const dh_tuning saved = g_tuning;
  g_tuning.detect_max_tiles = 64;
  int n1 = 0, n2 = 0, n3 = 0;
  rc = sweep(DH_KNOB_DETECT_TILE_TARGET, &dh_tuning::detect_tile_target,
             {1024, 2048, 4096, 8192, 16384, 32768}, run, reps, st, trials, trial_cap, &n1);
  auto rest = [&](int used) { return trials ? trials + (used < trial_cap ? used : trial_cap) : nullptr; };
  auto left = [&](int used) { return trial_cap > used ? trial_cap - used : 0; };
  if (rc == DH_OK)
    rc = sweep(DH_KNOB_DETECT_MAX_TILES, &dh_tuning::detect_max_tiles, {4, 8, 16, 32, 64}, run,
               reps, st, rest(n1), left(n1), &n2);
  if (rc == DH_OK)
    rc = sweep(DH_KNOB_DETECT_WG_PER_CU, &dh_tuning::detect_wg_per_cu, {0, 2, 4, 8, 16, 32},
               run, reps, st, rest(n1 + n2), left(n1 + n2), &n3);
  int n4 = 0, n5 = 0;
  if (rc == DH_OK)  // survivor queue always / never
    rc = sweep(DH_KNOB_DETECT_QUEUE_MIN_ITEMS, &dh_tuning::detect_queue_min_items, {0, 1 << 30},
               run, reps, st, rest(n1 + n2 + n3), left(n1 + n2 + n3), &n4);
  if (rc == DH_OK)  // the queue's binary32 screen off / on
    rc = sweep(DH_KNOB_DETECT_FP32_SCREEN, &dh_tuning::detect_fp32_screen, {0, 1}, run, reps, st,
               rest(n1 + n2 + n3 + n4), left(n1 + n2 + n3 + n4), &n5);
  if (rc) {
    g_tuning = saved;
    return rc;
  }
  if (ntrials) *ntrials = n1 + n2 + n3 + n4 + n5;
  return run();
}

}  // extern "C"
