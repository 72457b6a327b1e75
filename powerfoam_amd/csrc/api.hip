// Library-level entry points: version and the per-thread error channel.
#include <stdarg.h>
#include <stdio.h>

#include "pf_common.hpp"

namespace pf {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_check(hipError_t e, const char* what) {
  return set_error(PF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace pf

extern "C" {

int pf_version(void) { return 1; }

const char* pf_last_error(void) { return pf::g_err; }

}  // extern "C"
