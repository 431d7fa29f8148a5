// Per-thread last error message of the C ABI (host only).
#include "../../include/mtl4gpu.h"

#include <string>

namespace m4g {

static thread_local std::string tl_last_error;

void set_last_error(const std::string& msg) { tl_last_error = msg; }

}  // namespace m4g

extern "C" const char* m4g_last_error(void) { return m4g::tl_last_error.c_str(); }
