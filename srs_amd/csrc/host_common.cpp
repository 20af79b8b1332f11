// host_common.cpp -- error reporting and small shared helpers.
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "srs_internal.h"

namespace {
thread_local char g_err[1024] = "";
}

namespace srs {

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = 0; }

bool output_due(double t, double dt_last, double interval) {
  return std::floor(t / interval) > std::floor((t - dt_last) / interval);
}

}  // namespace srs

extern "C" {
const char* srs_last_error(void) { return g_err; }
const char* srs_version(void) { return "srs_amd 0.1 (gfx950)"; }
}
