// Error state and version of the C ABI (include/matvis_hip.h).
#include <stdarg.h>
#include <string.h>

#include "common.h"

namespace mv {

static thread_local char g_msg[512] = {0};
static thread_local int g_code = 0;

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  g_code = code;
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(MV_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return MV_OK;
}

}  // namespace mv

extern "C" int mv_last_error(char *buf, size_t len) {
  if (buf && len) {
    strncpy(buf, mv::g_msg, len - 1);
    buf[len - 1] = '\0';
  }
  return mv::g_code;
}

extern "C" int mv_abi_version(void) { return 3; }
