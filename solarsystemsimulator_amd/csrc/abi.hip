// Error reporting, library identification and launch timing for the C-ABI.
#include "ssim_internal.h"

#include <mutex>
#include <string>
#include <vector>

namespace ssim {

namespace {
thread_local char g_last_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

void clear_error() { g_last_error[0] = '\0'; }

// Launch timing: while enabled, the main kernel launches of the library are
// bracketed by two HIP events on their own stream; ssim_timing_report sums
// the elapsed times per kernel name.  Off by default (no events recorded).
namespace {
struct Timed {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_tmu;
std::vector<Timed> g_timed;
bool g_timing = false;

void drop_records() {
  for (Timed& t : g_timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  g_timed.clear();
}
}  // namespace

bool timing_on() { return g_timing; }

void* timing_start(hipStream_t st) {
  hipEvent_t a;
  if (hipEventCreate(&a) != hipSuccess) return nullptr;
  (void)hipEventRecord(a, st);
  return a;
}

void timing_stop(const char* name, void* start, hipStream_t st) {
  if (!start) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess) return;
  (void)hipEventRecord(b, st);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timed.push_back({name, static_cast<hipEvent_t>(start), b});
}

}  // namespace ssim

extern "C" {

const char* ssim_last_error(void) { return ssim::g_last_error; }

const char* ssim_version(void) { return "libssim 0.1 gfx950"; }

int ssim_timing(int enable) {
  std::lock_guard<std::mutex> lk(ssim::g_tmu);
  ssim::drop_records();
  ssim::g_timing = enable != 0;
  return SSIM_OK;
}

int ssim_timing_report(const char* kernel, double* total_ms, long long* launches) {
  SSIM_REQUIRE(kernel && total_ms && launches, "null pointer");
  std::lock_guard<std::mutex> lk(ssim::g_tmu);
  double tot = 0.0;
  long long cnt = 0;
  for (ssim::Timed& t : ssim::g_timed) {
    if (t.name != kernel) continue;
    SSIM_HIP(hipEventSynchronize(t.b));
    float ms = 0.f;
    SSIM_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return SSIM_OK;
}

int ssim_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

}  // extern "C"
