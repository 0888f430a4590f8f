// runtime.hpp — shared host-side plumbing of libdelta_hip (error state,
// device scratch buffers, launch helpers).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/delta_hip.h"

namespace dh {

// Records a message for dh_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...);

// Ensures dh_init() ran; returns DH_OK or an error code.
int ensure_init();

#define DH_HIP_CHECK(expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::dh::set_error(DH_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                             hipGetErrorString(e_), __FILE__, __LINE__);          \
  } while (0)

// A growable device buffer owned by the library (one per call site and
// device).  grow() keeps the old allocation when it is already big enough.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int grow(size_t need);
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

}  // namespace dh
