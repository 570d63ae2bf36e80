// Error reporting for the C ABI (thread-local last-error string).
#include "common.h"

namespace flux {

static thread_local char g_last_error[1024] = "no error";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace flux

extern "C" const char *flux_last_error(void) { return flux::g_last_error; }

extern "C" const char *flux_version(void) { return "fluxhip 0.1 gfx950"; }
