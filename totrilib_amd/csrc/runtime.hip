// Host-side plumbing shared by all entry points: per-thread error message,
// launch checking and device queries.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace totri_hip {

static thread_local char g_msg[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(TOTRI_EHIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return TOTRI_OK;
}

int device_cu_count() {
  static thread_local int cached_dev = -1;
  static thread_local int cached_cus = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev != cached_dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached_dev = dev;
    cached_cus = cus;
  }
  return cached_cus;
}

}  // namespace totri_hip

extern "C" const char* totri_last_error(void) { return totri_hip::g_msg; }

extern "C" int totri_abi_version(void) { return 101; }
