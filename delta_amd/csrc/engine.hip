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
                      long long* offs_out, dh_stats* stats, hipStream_t st);

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
  if (!s || !p) return set_error(DH_ERR_ARG, "dh_step: bad arguments");
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
                             g_step.counts.as<int>(), g_step.offs.as<long long>(), &stats, st);
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
  rc = dh_integrate(&s->bodies, g_step.force.as<double>(), g_step.torque.as<double>(),
                    g_step.touched.as<int>(), p->dt, p->gravity, s->x, s->y, s->z, s->refx,
                    s->refy, s->refz, s->geom.tri_offset, stream);
  if (rc) return rc;
  if (info) {
    info->candidate_pairs = np;
    info->contacts = total;
    info->triangle_checks = stats.triangle_checks;
  }
  return DH_OK;
}
