// Version and error reporting of the C ABI.
#include "aj_common.h"

#include <cstdio>

namespace aj {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  g_last_error = buf;
  return AJ_ERR_HIP;
}

}  // namespace aj

extern "C" const char* aj_version(void) { return "aerojax-amd 0.1.0 gfx950"; }

extern "C" const char* aj_last_error(void) { return aj::g_last_error.c_str(); }
