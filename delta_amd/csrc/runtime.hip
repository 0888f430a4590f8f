// runtime.hip — device selection, error reporting, memory and event helpers.
#include "runtime.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace dh {

static thread_local char g_err[512] = "";
static std::once_flag g_once;
static int g_init_status = -1;

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int ensure_init() {
  std::call_once(g_once, [] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
      g_init_status = set_error(DH_ERR_NO_DEVICE, "no GPU visible (%s)", hipGetErrorString(e));
      return;
    }
    e = hipSetDevice(0);
    if (e != hipSuccess) {
      g_init_status = set_error(DH_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
      return;
    }
    g_init_status = DH_OK;
  });
  if (g_init_status != DH_OK && g_err[0] == '\0')
    set_error(g_init_status, "initialisation failed");
  return g_init_status;
}

// While a captured step graph exists, its nodes hold the addresses of the
// library's scratch buffers: a buffer outgrown meanwhile is retired, not
// freed, until the last graph is destroyed.
static int g_graphs_alive = 0;
static std::vector<void*> g_retired;

void graphs_alive_add(int delta) {
  g_graphs_alive += delta;
  if (g_graphs_alive == 0) {
    for (void* p : g_retired) (void)hipFree(p);
    g_retired.clear();
  }
}

int DevBuf::grow(size_t need) {
  if (need <= bytes && ptr) return DH_OK;
  if (ptr && g_graphs_alive > 0)
    g_retired.push_back(ptr);
  else if (ptr)
    DH_HIP_CHECK(hipFree(ptr));
  ptr = nullptr;
  bytes = 0;
  size_t n = need < 256 ? 256 : need;
  DH_HIP_CHECK(hipMalloc(&ptr, n));
  bytes = n;
  return DH_OK;
}

}  // namespace dh

using namespace dh;

extern "C" {

int dh_init(void) { return ensure_init(); }

const char* dh_last_error(void) { return g_err; }

int dh_device_count(int* n) {
  if (!n) return set_error(DH_ERR_ARG, "dh_device_count: null");
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) *n = 0;
  return DH_OK;
}

int dh_malloc(void** p, size_t bytes) {
  if (!p) return set_error(DH_ERR_ARG, "dh_malloc: null");
  int rc = ensure_init();
  if (rc) return rc;
  DH_HIP_CHECK(hipMalloc(p, bytes ? bytes : 1));
  return DH_OK;
}

int dh_free(void* p) {
  if (!p) return DH_OK;
  DH_HIP_CHECK(hipFree(p));
  return DH_OK;
}

int dh_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  if (!bytes) return DH_OK;
  if (!dst || !src) return set_error(DH_ERR_ARG, "dh_memcpy_h2d: null");
  DH_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return DH_OK;
}

int dh_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  if (!bytes) return DH_OK;
  if (!dst || !src) return set_error(DH_ERR_ARG, "dh_memcpy_d2h: null");
  DH_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return DH_OK;
}

int dh_memset(void* p, int value, size_t bytes) {
  if (!bytes) return DH_OK;
  if (!p) return set_error(DH_ERR_ARG, "dh_memset: null");
  DH_HIP_CHECK(hipMemset(p, value, bytes));
  return DH_OK;
}

int dh_mem_info(size_t* free_bytes, size_t* total_bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  if (!free_bytes || !total_bytes) return set_error(DH_ERR_ARG, "dh_mem_info: null");
  DH_HIP_CHECK(hipMemGetInfo(free_bytes, total_bytes));
  return DH_OK;
}

int dh_synchronize(void) {
  int rc = ensure_init();
  if (rc) return rc;
  DH_HIP_CHECK(hipDeviceSynchronize());
  return DH_OK;
}

int dh_event_create(void** ev) {
  if (!ev) return set_error(DH_ERR_ARG, "dh_event_create: null");
  int rc = ensure_init();
  if (rc) return rc;
  hipEvent_t e;
  DH_HIP_CHECK(hipEventCreate(&e));
  *ev = e;
  return DH_OK;
}

int dh_event_destroy(void* ev) {
  if (ev) DH_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return DH_OK;
}

int dh_event_record(void* ev, void* stream) {
  DH_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream)));
  return DH_OK;
}

int dh_event_elapsed_ms(float* ms, void* start, void* stop) {
  DH_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  DH_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                   static_cast<hipEvent_t>(stop)));
  return DH_OK;
}

}  // extern "C"
