// Error reporting shared by every entry point of the library.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace tla {
static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}
}  // namespace tla

extern "C" const char* tla_last_error(void) { return tla::g_last_error; }

extern "C" int tla_abi_version(void) { return 1; }
