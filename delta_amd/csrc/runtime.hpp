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

// Step graphs alive (+1 on creation, -1 on destruction); see DevBuf::grow.
void graphs_alive_add(int delta);

// Launch parameters of the hot kernels (process-wide table of tuning.hip,
// set by dh_tuning_set and the autotuner).
const dh_tuning& tuning();

// Compute units of the current device (256 on MI355X).
int device_cus();

// Rows (triangles of A) per block that detection stages in LDS.  The
// broad phase's vertex pass keeps, for a particle i with more triangles than
// this, the box of each chunk c of 64 of its triangles at
// cbox[tri_offset[i] / 64 + i + c].
constexpr int kRowBlock = 320;

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// p[0..n) = v.  Used instead of hipMemsetAsync in work that may be captured
// into a HIP graph.
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, T v, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
inline void fill_async(T* p, T v, long long n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill<T>, grid_for(n, 256), 256, 0, st, p, v, n);
}

// Up to four regions cleared to zero by one launch (region y of the grid);
// sizes in 32-bit words.  Used instead of several fills in captured work.
struct ZeroRegions {
  unsigned* p[4];
  long long n[4];
};

__global__ __launch_bounds__(256) inline void k_zero_regions(ZeroRegions z) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const int r = blockIdx.y;
  if (i < z.n[r]) z.p[r][i] = 0u;
}

inline void zero_async(ZeroRegions z, hipStream_t st) {
  long long m = 0;
  for (int r = 0; r < 4; ++r) {
    if (!z.p[r]) z.n[r] = 0;
    m = z.n[r] > m ? z.n[r] : m;
  }
  if (m > 0) hipLaunchKernelGGL(k_zero_regions, dim3(grid_for(m, 256), 4), 256, 0, st, z);
}

}  // namespace dh
