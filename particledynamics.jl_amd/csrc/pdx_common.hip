// Error plumbing and version of libpdx.
#include "pdx_internal.hpp"

namespace pdx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

}  // namespace pdx

extern "C" int pdx_version(void) { return 100; }

extern "C" const char *pdx_last_error(void) { return pdx::g_last_error.c_str(); }
