// adaptive.hip — adaptive time-step control of the reference's Peano DEM
// runner on the device:
//   triggerParticleTooClose   multiscale/dem/mappings/Collision.cpp:188-299
//   TwoParticlesAreClose      State.cpp:122-135 (minimum of the positive
//                             proposals), merged like State.cpp:81-92
//   travel speed              MoveParticles.cpp:84-89 (fastest moved particle)
//   finishedTimeStep,         State.cpp:137-199 and clearAccumulatedData
//   adaptiveTimeStep          (State.cpp:26-35)
// Per step: k_adaptive_begin clears the reductions, k_adaptive_pairs runs
// before detection on the candidate pairs (positions and velocities at the
// start of the step, the step's dt), the step integrates with the state's dt
// (read from device memory, so a captured step follows it), then
// k_adaptive_travel and k_adaptive_update produce the next dt.
//
// Two choices where the reference is not well defined:
//  - MaxVelocityApproach: the reference stores vBA (negative) when -vBA beats
//    it (Collision.cpp:251-254), so its value depends on the order in which
//    its threads see the pairs; here it is the maximum of -vBA over the
//    approaching pairs and the previous step's travel speed, the quantity
//    its thread merge takes the maximum of (State.cpp:94-97).
//  - The state before the first update: StepIncrement = the increase factor
//    and _maxdt = +inf (the reference starts from unset values).
// Reductions are min / max / or, so the result does not depend on the order.
#include "runtime.hpp"

namespace dh {

__device__ __forceinline__ unsigned long long dbits(double v) {
  return static_cast<unsigned long long>(__double_as_longlong(v));
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
  return __longlong_as_double(static_cast<long long>(b));
}

__global__ void k_adaptive_begin(dh_adaptive_state* s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    s->scratch[0] = dbits(INFINITY);             // close: minimum of positive proposals
    s->scratch[1] = dbits(s->max_vel_travel);    // approach: starts at the last travel speed
    s->scratch[2] = 0;                           // travel speed of this step (bits, >= 0)
    s->scratch[3] = 0;                           // separate
  }
}

// Collision.cpp:188-299 for candidate pair k (A = pairs[2k], B = pairs[2k+1]).
__global__ __launch_bounds__(256) void k_adaptive_pairs(const int* __restrict__ pairs, int npairs,
                                                        const long long* __restrict__ np_dev,
                                                        const double* __restrict__ pos,
                                                        const double* __restrict__ linvel,
                                                        const double* __restrict__ radius,
                                                        const double* __restrict__ eps,
                                                        dh_adaptive_state* __restrict__ s) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  bool active = k < npairs && !(np_dev && k >= *np_dev);
  double close = INFINITY, approach = 0.0;
  bool sep = false;
  if (active) {
    const int a = pairs[2 * k], b = pairs[2 * k + 1];
    const double increments = s->increase;
    const double dt = s->dt;
    double pdt = dt * increments;
    if (pdt >= s->max_dt) pdt = s->max_dt;
    double cA[3], cB[3], dAB[3], pdAB[3], vA[3], vB[3];
    for (int q = 0; q < 3; ++q) {
      vA[q] = linvel[3 * a + q];
      vB[q] = linvel[3 * b + q];
      cA[q] = pos[3 * a + q] + pdt * vA[q];
      cB[q] = pos[3 * b + q] + pdt * vB[q];
    }
    for (int q = 0; q < 3; ++q) {
      dAB[q] = pos[3 * b + q] - pos[3 * a + q];
      pdAB[q] = cB[q] - cA[q];
    }
    const double d = sqrt((dAB[0] * dAB[0]) + (dAB[1] * dAB[1]) + (dAB[2] * dAB[2]));
    const double pd = sqrt((pdAB[0] * pdAB[0]) + (pdAB[1] * pdAB[1]) + (pdAB[2] * pdAB[2]));
    const double vBA = ((vB[0] * dAB[0]) + (vB[1] * dAB[1]) + (vB[2] * dAB[2])) -
                       ((vA[0] * dAB[0]) + (vA[1] * dAB[1]) + (vA[2] * dAB[2]));
    const double pvBA = ((vB[0] * pdAB[0]) + (vB[1] * pdAB[1]) + (vB[2] * pdAB[2])) -
                        ((vA[0] * pdAB[0]) + (vA[1] * pdAB[1]) + (vA[2] * pdAB[2]));
    if (vBA > 0 && pvBA > 0) {
      sep = true;
    } else {
      approach = -vBA;
      // haloDiameter = (diameter + 2 epsilon) * 1.1 (Vertex.cpp:113)
      const double rA = ((2.0 * radius[a] + eps[a] * 2) * 1.1) / 2.0;
      const double rB = ((2.0 * radius[b] + eps[b] * 2) * 1.1) / 2.0;
      const double predictedMinHaloDistance = (pd - rA - rB);
      const double pdistancePerStep = predictedMinHaloDistance / pdt;
      if (-pvBA >= pdistancePerStep) {
        const double rrA = (2.0 * radius[a]) / 2.0, rrB = (2.0 * radius[b]) / 2.0;
        double remainingDistance = (d - rrA - rrB - eps[a] - eps[b]);
        if (remainingDistance < 0.0) remainingDistance = remainingDistance * -1.1;
        double localMaxdt = remainingDistance * 0.82;
        if (localMaxdt < 0.0) localMaxdt = -1 * localMaxdt;
        if (localMaxdt > 0.0) close = localMaxdt;
      }
    }
  }
  // wave reductions, then one atomic per wave
  // every lane takes part in each shuffle (no short-circuit around one)
  for (int o = 32; o > 0; o >>= 1) {
    close = fmin(close, __shfl_xor(close, o));
    approach = fmax(approach, __shfl_xor(approach, o));
    const int other = __shfl_xor(static_cast<int>(sep), o);
    sep = sep || other != 0;
  }
  if ((threadIdx.x & 63) == 0) {
    if (close < INFINITY) atomicMin(&s->scratch[0], dbits(close));
    if (approach > 0.0) atomicMax(&s->scratch[1], dbits(approach));
    if (sep) atomicOr(&s->scratch[3], 1ull);
  }
}

// MoveParticles.cpp:84-89: speed of every moved particle after the update.
__global__ __launch_bounds__(256) void k_adaptive_travel(const double* __restrict__ linvel,
                                                         const int* __restrict__ obstacle, int n,
                                                         dh_adaptive_state* __restrict__ s) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (i < n && !obstacle[i]) {
    const double* u = linvel + 3 * i;
    v = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(&s->scratch[2], dbits(v));
}

// finishedTimeStep, adaptiveTimeStep and clearAccumulatedData.
__global__ void k_adaptive_update(dh_adaptive_state* s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  s->time = s->time + s->dt;
  const double close_min = bitsd(s->scratch[0]);
  s->close = close_min < INFINITY ? close_min : 0.0;
  s->max_vel_approach = bitsd(s->scratch[1]);
  s->separate = s->scratch[3] ? 1 : 0;
  const double increaseFactor = s->increase;
  const double mindt = s->min_dt;
  s->max_dt = s->mesh_width / (2.0 * increaseFactor * s->max_vel_approach) * 0.80;
  bool clamp = true;
  if (s->close > 0.0) {
    if (s->dt >= mindt) {
      if (!(s->close > increaseFactor * s->dt)) s->dt = s->close;
      clamp = false;  // State.cpp:176 returns before the upper bound
    } else if (s->dt < mindt) {
      s->dt = mindt;
    }
  } else if (s->step > 2) {
    s->dt = increaseFactor * s->dt;
  }
  if (clamp && s->dt > s->max_dt) s->dt = s->max_dt;
  s->max_vel_travel = bitsd(s->scratch[2]);
  s->step = s->step + 1;
}

int adaptive_begin(dh_adaptive_state* s, hipStream_t st) {
  hipLaunchKernelGGL(k_adaptive_begin, 1, 64, 0, st, s);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

int adaptive_pairs(const int* pairs, int npairs, const long long* np_dev, const double* pos,
                   const double* linvel, const double* radius, const double* eps,
                   dh_adaptive_state* s, hipStream_t st) {
  if (npairs > 0) {
    hipLaunchKernelGGL(k_adaptive_pairs, grid_for(npairs, 256), 256, 0, st, pairs, npairs, np_dev,
                       pos, linvel, radius, eps, s);
    DH_HIP_CHECK(hipGetLastError());
  }
  return DH_OK;
}

int adaptive_finish(const double* linvel, const int* obstacle, int n, dh_adaptive_state* s,
                    hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_adaptive_travel, grid_for(n, 256), 256, 0, st, linvel, obstacle, n, s);
  hipLaunchKernelGGL(k_adaptive_update, 1, 64, 0, st, s);
  DH_HIP_CHECK(hipGetLastError());
  return DH_OK;
}

}  // namespace dh

using namespace dh;

extern "C" int dh_adaptive_step(const int* pairs_dev, int npairs, const dh_bodies* bodies,
                                const double* radius_dev, const double* eps_dev,
                                dh_adaptive_state* state_dev, int phase, void* stream) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!bodies || !state_dev || npairs < 0 || (npairs > 0 && (!pairs_dev || !radius_dev || !eps_dev)))
    return set_error(DH_ERR_ARG, "dh_adaptive_step: bad arguments");
  const hipStream_t st = as_stream(stream);
  if (phase == 0) {
    if ((rc = adaptive_begin(state_dev, st))) return rc;
    return adaptive_pairs(pairs_dev, npairs, nullptr, bodies->pos, bodies->linvel, radius_dev,
                          eps_dev, state_dev, st);
  }
  if (phase == 1) return adaptive_finish(bodies->linvel, bodies->obstacle, bodies->n, state_dev, st);
  return set_error(DH_ERR_ARG, "dh_adaptive_step: phase %d (0 pairs, 1 finish)", phase);
}
