// mixed.hip — FP32-screened / learned unique-pair paths (filled in below).
#include "runtime.hpp"

namespace dh {
int tri_pairs_extra(int method, const double*, const float*, int64_t, double, double*, double*,
                    double*, uint8_t*, hipStream_t) {
  return set_error(DH_ERR_UNSUPPORTED, "dh_tri_pairs: method %d not built yet", method);
}
}  // namespace dh

extern "C" int dh_tri_pairs_counters(dh_pair_counters* out) {
  if (!out) return ::dh::set_error(DH_ERR_ARG, "null");
  *out = dh_pair_counters{0, 0, 0, 0};
  return DH_OK;
}
