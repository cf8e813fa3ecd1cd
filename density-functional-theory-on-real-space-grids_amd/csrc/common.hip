// Error recording shared by every entry point.
#include "common.h"

namespace cdft {

static thread_local char g_message[1024] = "";

const char* last_error() { return g_message; }

int fail(int code, const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_message, sizeof(g_message), fmt, args);
  va_end(args);
  return code;
}

int check_launch(const char* what) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    return fail(static_cast<int>(err), "%s: kernel launch failed: %s", what, hipGetErrorString(err));
  }
  return CDFT_OK;
}

}  // namespace cdft

extern "C" const char* cdft_last_error(void) { return cdft::last_error(); }

extern "C" int cdft_abi_version(void) { return 1; }
