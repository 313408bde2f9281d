// Error state and version of the C ABI.
#include "gsevt_api.h"

namespace gsevt {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace gsevt

extern "C" {

// Message of the last failing call on this thread ("" if none).
const char* gsevt_last_error(void) { return gsevt::g_err; }

void gsevt_clear_error(void) { gsevt::g_err[0] = '\0'; }

int gsevt_abi_version(void) { return 1; }

}  // extern "C"
