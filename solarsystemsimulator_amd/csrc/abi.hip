// Error reporting and library identification for the C-ABI.
#include "ssim_internal.h"

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

}  // namespace ssim

extern "C" {

const char* ssim_last_error(void) { return ssim::g_last_error; }

const char* ssim_version(void) { return "libssim 0.1 gfx950"; }

int ssim_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

}  // extern "C"
