// mixed.hip — dispatch of the non-exact unique-pair paths.
#include "runtime.hpp"

namespace dh {
int tri_pairs_learned(const double* soa, int64_t n, double* P, double* Q, double* d,
                      uint8_t* flags, hipStream_t st);
int learned_counters(unsigned long long out[3]);

int tri_pairs_extra(int method, const double* soa, const float*, int64_t n, double, double* P,
                    double* Q, double* d, uint8_t* flags, hipStream_t st) {
  if (method == DH_PAIR_LEARNED) return tri_pairs_learned(soa, n, P, Q, d, flags, st);
  return set_error(DH_ERR_UNSUPPORTED, "dh_tri_pairs: method %d not available", method);
}
}  // namespace dh

extern "C" int dh_tri_pairs_counters(dh_pair_counters* out) {
  if (!out) return ::dh::set_error(DH_ERR_ARG, "null");
  unsigned long long c[3];
  int rc = ::dh::learned_counters(c);
  if (rc) return rc;
  out->screened = 0;
  out->certified = static_cast<int64_t>(c[0]);
  out->features_fp64 = static_cast<int64_t>(c[1]);
  out->rechecked = static_cast<int64_t>(c[2]);
  return DH_OK;
}
