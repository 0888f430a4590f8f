// engine.hip — one explicit DEM step over a device-resident scene, in the
// order of Engine::iterate (Engine.cpp:155-163): broad phase, narrow phase on
// the current geometry, contact forces, velocity/angular update, gravity,
// position, rotation and vertex replicas.
#include "runtime.hpp"

namespace dh {

int broadphase_impl(const dh_geometry* geom, const int* obstacle_dev, double cell, int* pairs,
                    long long cap, long long* npairs, hipStream_t st);
int detect_pairs_impl(int method, const dh_geometry* geom, const int* pairs, int npairs,
                      dh_contact* out, long long cap, long long* n_total, int* pair_count_user,
                      long long* offs_out, dh_stats* stats, hipStream_t st, const double* pbox,
                      const double* bbox, const double* psph);
const double* broadphase_boxes();
const double* broadphase_block_boxes();
const double* broadphase_spheres();

int broadphase_graph(const dh_geometry* geom, const int* obstacle_dev, double cell, int* pairs,
                     long long cap, const long long** np_dev, hipStream_t st);
int broadphase_reserve(int n, int ntri);
int detect_pairs_graph(int method, const dh_geometry* geom, const int* pairs, int pair_cap,
                       const long long* np_dev, dh_contact* out, long long contact_cap,
                       int* counts, long long* offs, hipStream_t st, const double* pbox,
                       const double* bbox, const double* psph);
int detect_reserve(int pair_cap, long long contact_cap, int method);
int detect_prepare(const dh_geometry* geom, hipStream_t st);
const unsigned long long* detect_counter_dev();
const int* detect_error_dev();
const unsigned long long* detect_stats_dev();
int forces_reserve(int n, int npairs);
int contacts_forces_impl(const dh_bodies* bodies, const int* pairs_dev, int npairs,
                         const long long* np_dev, const dh_contact* contacts_dev,
                         long long contact_cap, const long long* pair_offset_dev,
                         const dh_force_params* fp, double* force_dev, double* torque_dev,
                         int* touched_dev, hipStream_t st);
int scan_counts(const int* counts, int n, long long* offs, hipStream_t st);
int integrate_impl(const dh_bodies* bodies, const double* force_dev, const double* torque_dev,
                   const int* touched_dev, double dt, const double* dt_dev,
                   const double gravity[3], double* x_dev, double* y_dev, double* z_dev,
                   const double* refx_dev, const double* refy_dev, const double* refz_dev,
                   const int* tri_offset_dev, hipStream_t st);
int adaptive_begin(dh_adaptive_state* s, hipStream_t st);
int adaptive_pairs(const int* pairs, int npairs, const long long* np_dev, const double* pos,
                   const double* linvel, const double* radius, const double* eps,
                   dh_adaptive_state* s, hipStream_t st);
int adaptive_finish(const double* linvel, const int* obstacle, int n, dh_adaptive_state* s,
                    hipStream_t st);

struct StepScratch {
  DevBuf pairs, contacts, counts, offs, force, torque, touched;
  long long pair_cap = 0, contact_cap = 0;
};
static StepScratch g_step;

}  // namespace dh

using namespace dh;

extern "C" int dh_step(dh_scene* s, const dh_step_params* p, dh_step_info* info, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!s || !p || (p->adaptive && !p->radius))
    return set_error(DH_ERR_ARG, "dh_step: bad arguments");
  const hipStream_t st = as_stream(stream);
  const int n = s->geom.nparticles;
  if (g_step.pair_cap == 0) g_step.pair_cap = 16LL * (n + 16);
  if (g_step.contact_cap == 0) g_step.contact_cap = 1 << 16;
  long long np = 0;
  for (;;) {
    if ((rc = g_step.pairs.grow(sizeof(int) * 2 * g_step.pair_cap))) return rc;
    rc = broadphase_impl(&s->geom, s->bodies.obstacle, p->cell, g_step.pairs.as<int>(),
                         g_step.pair_cap, &np, st);
    if (rc != DH_ERR_CAPACITY) break;
    g_step.pair_cap = 2 * np;
  }
  if (rc) return rc;
  if (p->adaptive) {
    // triggerParticleTooClose on the candidates, at the start of the step
    if ((rc = adaptive_begin(p->adaptive, st))) return rc;
    if ((rc = adaptive_pairs(g_step.pairs.as<int>(), static_cast<int>(np), nullptr, s->bodies.pos,
                             s->bodies.linvel, p->radius, s->geom.eps, p->adaptive, st)))
      return rc;
  }
  if ((rc = g_step.counts.grow(sizeof(int) * (np + 1)))) return rc;
  if ((rc = g_step.offs.grow(sizeof(long long) * (np + 2)))) return rc;
  long long total = 0;
  dh_stats stats{};
  if (np > 0) {
    for (;;) {
      if ((rc = g_step.contacts.grow(sizeof(dh_contact) * g_step.contact_cap))) return rc;
      stats = dh_stats{};
      rc = detect_pairs_impl(p->method, &s->geom, g_step.pairs.as<int>(), static_cast<int>(np),
                             g_step.contacts.as<dh_contact>(), g_step.contact_cap, &total,
                             g_step.counts.as<int>(), g_step.offs.as<long long>(), &stats, st,
                             broadphase_boxes(), broadphase_block_boxes(), broadphase_spheres());
      if (rc != DH_ERR_CAPACITY) break;
      g_step.contact_cap = 2 * total;
    }
    if (rc) return rc;
  }
  if ((rc = g_step.force.grow(sizeof(double) * 3 * (n + 1)))) return rc;
  if ((rc = g_step.torque.grow(sizeof(double) * 3 * (n + 1)))) return rc;
  if ((rc = g_step.touched.grow(sizeof(int) * (n + 1)))) return rc;
  if (np > 0) {
    rc = dh_contacts_forces(&s->bodies, g_step.pairs.as<int>(), static_cast<int>(np),
                            g_step.contacts.as<dh_contact>(),
                            reinterpret_cast<const int64_t*>(g_step.offs.ptr), &p->forces,
                            g_step.force.as<double>(), g_step.torque.as<double>(),
                            g_step.touched.as<int>(), stream);
    if (rc) return rc;
  } else {
    DH_HIP_CHECK(hipMemsetAsync(g_step.touched.ptr, 0, sizeof(int) * n, st));
  }
  rc = integrate_impl(&s->bodies, g_step.force.as<double>(), g_step.torque.as<double>(),
                      g_step.touched.as<int>(), p->dt, p->adaptive ? &p->adaptive->dt : nullptr,
                      p->gravity, s->x, s->y, s->z, s->refx, s->refy, s->refz,
                      s->geom.tri_offset, st);
  if (rc) return rc;
  if (p->adaptive &&
      (rc = adaptive_finish(s->bodies.linvel, s->bodies.obstacle, n, p->adaptive, st)))
    return rc;
  if (info) {
    info->candidate_pairs = np;
    info->contacts = total;
    info->triangle_checks = stats.triangle_checks;
    info->pairs_culled = stats.pairs_culled;
  }
  return DH_OK;
}

// ------------------------------------------------------ captured step graph
namespace dh {

// Status of the last replay: candidate pairs, contacts found, error flag,
// triangle checks (written by the graph's last node), and in out[4] the
// overflow bits of every replay since the status was last read (1 pairs,
// 2 contacts, 4 detection error).
__global__ void k_step_status(const long long* __restrict__ np,
                              const unsigned long long* __restrict__ contacts,
                              const int* __restrict__ err,
                              const unsigned long long* __restrict__ stats, long long pair_cap,
                              long long contact_cap, long long* __restrict__ out) {
  if (threadIdx.x == 0) {
    const long long p = *np, c = static_cast<long long>(*contacts);
    const int e = *err;
    out[0] = p;
    out[1] = c;
    out[2] = e;
    out[3] = static_cast<long long>(stats[20]);  // triangle checks
    out[4] |= (p > pair_cap ? 1 : 0) | (c > contact_cap ? 2 : 0) | (e ? 4 : 0);
    out[5] = static_cast<long long>(stats[23]);  // pairs culled by the box test
  }
}

}  // namespace dh

struct dh_step_graph {
  dh_scene scene;
  dh_step_params params;
  long long pair_cap, contact_cap;
  DevBuf pairs, counts, offs, contacts, force, torque, touched, status;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int record_step(dh_step_graph* g, hipStream_t st) {
  int rc;
  const dh_scene* s = &g->scene;
  const long long* np_dev = nullptr;
  if ((rc = broadphase_graph(&s->geom, s->bodies.obstacle, g->params.cell, g->pairs.as<int>(),
                             g->pair_cap, &np_dev, st)))
    return rc;
  dh_adaptive_state* ad = g->params.adaptive;
  if (ad) {
    if ((rc = adaptive_begin(ad, st))) return rc;
    if ((rc = adaptive_pairs(g->pairs.as<int>(), static_cast<int>(g->pair_cap), np_dev,
                             s->bodies.pos, s->bodies.linvel, g->params.radius, s->geom.eps, ad,
                             st)))
      return rc;
  }
  if ((rc = detect_pairs_graph(g->params.method, &s->geom, g->pairs.as<int>(),
                               static_cast<int>(g->pair_cap), np_dev,
                               g->contacts.as<dh_contact>(), g->contact_cap, g->counts.as<int>(),
                               g->offs.as<long long>(), st, broadphase_boxes(),
                               broadphase_block_boxes(), broadphase_spheres())))
    return rc;
  if ((rc = contacts_forces_impl(&s->bodies, g->pairs.as<int>(), static_cast<int>(g->pair_cap),
                                 np_dev, g->contacts.as<dh_contact>(), g->contact_cap,
                                 g->offs.as<long long>(),
                                 &g->params.forces, g->force.as<double>(),
                                 g->torque.as<double>(), g->touched.as<int>(), st)))
    return rc;
  if ((rc = integrate_impl(&s->bodies, g->force.as<double>(), g->torque.as<double>(),
                           g->touched.as<int>(), g->params.dt, ad ? &ad->dt : nullptr,
                           g->params.gravity, s->x, s->y, s->z, s->refx, s->refy, s->refz,
                           s->geom.tri_offset, st)))
    return rc;
  if (ad && (rc = adaptive_finish(s->bodies.linvel, s->bodies.obstacle, s->bodies.n, ad, st)))
    return rc;
  hipLaunchKernelGGL(k_step_status, 1, 64, 0, st, np_dev, detect_counter_dev(), detect_error_dev(),
                     detect_stats_dev(), g->pair_cap, g->contact_cap, g->status.as<long long>());
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

static void destroy_graph(dh_step_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  for (DevBuf* b : {&g->pairs, &g->counts, &g->offs, &g->contacts, &g->force, &g->torque,
                    &g->touched, &g->status})
    if (b->ptr) (void)hipFree(b->ptr);
  delete g;
}

extern "C" {

int dh_step_graph_create(const dh_scene* scene, const dh_step_params* params, int64_t pair_cap,
                         int64_t contact_cap, dh_step_graph** out) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!scene || !params || !out || pair_cap <= 0 || contact_cap <= 0 ||
      (params->adaptive && !params->radius) ||
      pair_cap > 0x7fffffffLL || scene->geom.nparticles < 1)
    return set_error(DH_ERR_ARG, "dh_step_graph_create: bad arguments");
  if (params->method != DH_BF && params->method != DH_PENALTY && params->method != DH_BF_LEARNED)
    return set_error(DH_ERR_UNSUPPORTED,
                     "dh_step_graph_create: method %d (DH_BF, DH_PENALTY or DH_BF_LEARNED)",
                     params->method);
  *out = nullptr;
  dh_step_graph* g = new dh_step_graph();
  g->scene = *scene;
  g->params = *params;
  g->pair_cap = pair_cap;
  g->contact_cap = contact_cap;
  const int n = scene->geom.nparticles;
  // every allocation before the capture
  if ((rc = g->pairs.grow(sizeof(int) * 2 * pair_cap)) ||
      (rc = g->counts.grow(sizeof(int) * (pair_cap + 1))) ||
      (rc = g->offs.grow(sizeof(long long) * (pair_cap + 2))) ||
      (rc = g->contacts.grow(sizeof(dh_contact) * contact_cap)) ||
      (rc = g->force.grow(sizeof(double) * 3 * (n + 1))) ||
      (rc = g->torque.grow(sizeof(double) * 3 * (n + 1))) ||
      (rc = g->touched.grow(sizeof(int) * (n + 1))) ||
      (rc = g->status.grow(sizeof(long long) * 6)) || (rc = broadphase_reserve(n, scene->geom.ntri_total)) ||
      (rc = detect_reserve(static_cast<int>(pair_cap > n ? pair_cap : n), contact_cap,
                           params->method)) ||
      (rc = detect_reserve(static_cast<int>(pair_cap), contact_cap, params->method)) ||
      (rc = forces_reserve(n, static_cast<int>(pair_cap)))) {
    destroy_graph(g);
    return rc;
  }
  // a blocking stream: replays on it start after earlier work on the null
  // stream, whose entry points (dh_step, dh_detect_pairs, ...) share the
  // library's scratch buffers with the graph
  hipError_t e = hipStreamCreate(&g->stream);
  if (e == hipSuccess) e = hipMemsetAsync(g->status.ptr, 0, sizeof(long long) * 6, g->stream);
  // what detection learns of the geometry with a read-back, before the capture
  if (e == hipSuccess) detect_prepare(&scene->geom, g->stream);
  if (e == hipSuccess) e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    destroy_graph(g);
    return set_error(DH_ERR_HIP, "step graph capture: %s", hipGetErrorString(e));
  }
  rc = record_step(g, g->stream);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(g->stream, &graph);
  g->graph = graph;
  if (rc != DH_OK) {
    destroy_graph(g);
    return rc;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    destroy_graph(g);
    return set_error(DH_ERR_HIP, "step graph instantiate: %s", hipGetErrorString(e));
  }
  graphs_alive_add(1);
  *out = g;
  return DH_OK;
}

}  // extern "C"

namespace dh {

// Every replay shares the library's scratch with every other graph's
// replays: each launch waits for the previous launch of any graph (an event
// recorded after it), on whatever stream either ran.
static hipEvent_t g_replay_done = nullptr;
static bool g_replay_recorded = false;

}  // namespace dh

extern "C" {

int dh_step_graph_launch(dh_step_graph* g, void* stream) {
  if (!g || !g->exec) return set_error(DH_ERR_ARG, "dh_step_graph_launch: no graph");
  const hipStream_t st = stream ? as_stream(stream) : g->stream;
  if (stream) {
    // a non-blocking stream would not wait for the eager entry points' work
    // on the null stream, which uses the same scratch
    unsigned flags = 0;
    DH_HIP_CHECK(hipStreamGetFlags(st, &flags));
    if (flags & hipStreamNonBlocking)
      return set_error(DH_ERR_ARG, "dh_step_graph_launch: non-blocking stream");
  }
  if (!g_replay_done)
    DH_HIP_CHECK(hipEventCreateWithFlags(&g_replay_done, hipEventDisableTiming));
  if (g_replay_recorded) DH_HIP_CHECK(hipStreamWaitEvent(st, g_replay_done, 0));
  DH_HIP_CHECK(hipGraphLaunch(g->exec, st));
  DH_HIP_CHECK(hipEventRecord(g_replay_done, st));
  g_replay_recorded = true;
  return DH_OK;
}

int dh_step_graph_status(dh_step_graph* g, void* stream, dh_step_info* info, int* overflow) {
  if (!g || !info || !overflow) return set_error(DH_ERR_ARG, "dh_step_graph_status: bad arguments");
  long long h[6];
  const hipStream_t st = stream ? as_stream(stream) : g->stream;
  DH_HIP_CHECK(hipMemcpyAsync(h, g->status.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
  // the sticky overflow word restarts at zero (outside the captured graph)
  DH_HIP_CHECK(hipMemsetAsync(g->status.as<long long>() + 4, 0, sizeof(long long), st));
  DH_HIP_CHECK(hipStreamSynchronize(st));
  info->candidate_pairs = h[0];
  info->contacts = h[1];
  info->triangle_checks = h[3];
  info->pairs_culled = h[5];
  *overflow = static_cast<int>(h[4]);
  return DH_OK;
}

int dh_step_graph_destroy(dh_step_graph* g) {
  if (!g) return DH_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  destroy_graph(g);
  graphs_alive_add(-1);
  return DH_OK;
}

}  // extern "C"
