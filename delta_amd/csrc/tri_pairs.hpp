// tri_pairs.hpp — parameters shared by the unique-pair kernels.
#pragma once
#include <cstdint>

namespace dh {

struct GenParams {
  double box;          // vertices uniform in [0, box)^3
  double sep_min;      // separation log-uniform in [sep_min, sep_max]
  double sep_max;
  double frac_parallel;
  double frac_intersect;
  double max_angle;    // near-parallel edge angle bound (rad)
  int bisection_steps;
};

}  // namespace dh
