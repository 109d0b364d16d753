// Error reporting and small C-ABI helpers of librgp.
#include <stdarg.h>
#include <stdio.h>

#include "rgp_internal.h"

static thread_local char g_err[512] = "";

// The host packs cameras as 32 float32 words; arrays of them have that stride.
static_assert(sizeof(rgp_camera) == 128, "rgp_camera is 32 words");

namespace rgp {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* who) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(RGP_ERR_LAUNCH, "%s: %s", who, hipGetErrorString(e));
  return RGP_OK;
}

}  // namespace rgp

extern "C" const char* rgp_last_error(void) { return g_err; }

extern "C" int rgp_version(void) { return 1; }
