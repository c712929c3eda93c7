// runtime.cpp -- error state, streams, launch timers and memory entry points.
#include "klrtm_internal.h"
#include <cstring>
#include <map>
#include <mutex>

namespace klrtm {

static thread_local std::string g_err;
void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
const char *get_error() { return g_err.c_str(); }

// ---- stream capture: launches are recorded into a graph, so the launch timers stay off ----
static thread_local bool t_capturing = false;
void set_capturing(bool on) { t_capturing = on; }
bool capturing() { return t_capturing; }

// ---- streams ------------------------------------------------------------------
static thread_local hipStream_t t_user_stream = nullptr;
static thread_local bool t_user_stream_set = false;
static std::mutex g_stream_mu;
static std::map<int, hipStream_t> g_dev_stream;

hipStream_t stream() {
  if(t_user_stream_set) return t_user_stream;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_stream_mu);
  auto it = g_dev_stream.find(dev);
  if(it != g_dev_stream.end()) return it->second;
  hipStream_t s = nullptr;
  if(hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  g_dev_stream[dev] = s;
  return s;
}

// ---- launch timers ----------------------------------------------------------------
struct TimerStat {
  double ms = 0;
  long count = 0;
};
static std::mutex g_tmu;
static bool g_timer_on = false;
static std::map<std::string, TimerStat> g_stats;
struct Pending {
  std::string name;
  hipEvent_t e0, e1;
};
static std::vector<Pending> g_pending;
static std::vector<hipEvent_t> g_pool;

static hipEvent_t get_event() {
  if(!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if(hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
static void drain_pending() {
  for(auto &p : g_pending) {
    float ms = 0;
    if(hipEventSynchronize(p.e1) == hipSuccess && hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
      auto &st = g_stats[p.name];
      st.ms += ms;
      st.count += 1;
    }
    g_pool.push_back(p.e0);
    g_pool.push_back(p.e1);
  }
  g_pending.clear();
}

LaunchTimer::LaunchTimer(const char *kernel_name, hipStream_t s) : name(kernel_name), st(s ? s : stream()) {
  if(t_capturing) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if(!g_timer_on) return;
  if(g_pending.size() > 4096) drain_pending();
  e0 = get_event();
  e1 = get_event();
  if(e0) (void)hipEventRecord(e0, st);
}
LaunchTimer::~LaunchTimer() {
  if(!e0 || !e1) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  (void)hipEventRecord(e1, st);
  g_pending.push_back({name, e0, e1});
}

int check_launch(const char *name) {
  hipError_t e = hipGetLastError();
  if(e != hipSuccess) {
    set_error("launch of %s failed: %s", name, hipGetErrorString(e));
    return KLRTM_EDEVICE;
  }
  return KLRTM_OK;
}

}  // namespace klrtm

using namespace klrtm;

extern "C" {

const char *klrtm_last_error(void) { return get_error(); }

int klrtm_version(int *major, int *minor) {
  if(major) *major = 0;
  if(minor) *minor = 1;
  return KLRTM_OK;
}

int klrtm_device_count(int *n) {
  KL_REQUIRE(n, KLRTM_EINVAL, "null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if(e != hipSuccess) c = 0;
  *n = c;
  return KLRTM_OK;
}
int klrtm_set_device(int device) {
  KL_HIP(hipSetDevice(device));
  return KLRTM_OK;
}
int klrtm_set_stream(void *s) {
  t_user_stream = (hipStream_t)s;
  t_user_stream_set = s != nullptr;
  return KLRTM_OK;
}
int klrtm_stream_create(void **s) {
  KL_REQUIRE(s, KLRTM_EINVAL, "null stream output");
  hipStream_t h = nullptr;
  KL_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));  // default priority
  *s = (void *)h;
  return KLRTM_OK;
}
int klrtm_stream_destroy(void *s) {
  if(s) KL_HIP(hipStreamDestroy((hipStream_t)s));
  return KLRTM_OK;
}
int klrtm_stream_synchronize(void *s) {
  KL_HIP(hipStreamSynchronize(s ? (hipStream_t)s : stream()));
  return KLRTM_OK;
}
int klrtm_malloc(void **ptr, size_t bytes) {
  KL_REQUIRE(ptr, KLRTM_EINVAL, "null pointer output");
  *ptr = nullptr;
  if(bytes == 0) return KLRTM_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if(e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return KLRTM_ENOMEM;
  }
  return KLRTM_OK;
}
int klrtm_free(void *ptr) {
  if(ptr) KL_HIP(hipFree(ptr));
  return KLRTM_OK;
}
int klrtm_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  if(!bytes) return KLRTM_OK;
  KL_REQUIRE(dst && src, KLRTM_EINVAL, "null pointer");
  KL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  KL_HIP(hipStreamSynchronize(stream()));
  return KLRTM_OK;
}
int klrtm_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if(!bytes) return KLRTM_OK;
  KL_REQUIRE(dst && src, KLRTM_EINVAL, "null pointer");
  KL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  KL_HIP(hipStreamSynchronize(stream()));
  return KLRTM_OK;
}
int klrtm_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  if(!bytes) return KLRTM_OK;
  KL_REQUIRE(dst && src, KLRTM_EINVAL, "null pointer");
  KL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  return KLRTM_OK;
}
int klrtm_memset(void *dst, int value, size_t bytes) {
  if(!bytes) return KLRTM_OK;
  KL_REQUIRE(dst, KLRTM_EINVAL, "null pointer");
  KL_HIP(hipMemsetAsync(dst, value, bytes, stream()));
  return KLRTM_OK;
}
int klrtm_synchronize(void) {
  KL_HIP(hipStreamSynchronize(stream()));
  return KLRTM_OK;
}

int klrtm_timer_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timer_on = on != 0;
  return KLRTM_OK;
}
int klrtm_timer_read(const char *kernel_name, double *total_ms, long *count) {
  KL_REQUIRE(kernel_name, KLRTM_EINVAL, "null kernel name");
  std::lock_guard<std::mutex> lk(g_tmu);
  drain_pending();
  auto it = g_stats.find(kernel_name);
  if(total_ms) *total_ms = it == g_stats.end() ? 0.0 : it->second.ms;
  if(count) *count = it == g_stats.end() ? 0 : it->second.count;
  return KLRTM_OK;
}
int klrtm_timer_reset(void) {
  std::lock_guard<std::mutex> lk(g_tmu);
  drain_pending();
  g_stats.clear();
  return KLRTM_OK;
}

}  // extern "C"
