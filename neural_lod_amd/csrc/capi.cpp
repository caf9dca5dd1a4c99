// Library-wide C ABI entry points: error reporting and version.
#include <string>

#include "nlod_common.h"

namespace nlod {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
}  // namespace nlod

extern "C" {
const char* nlod_last_error(void) { return nlod::g_last_error.c_str(); }
const char* nlod_version(void) { return "nlod-mi355x 0.1.0 (gfx950)"; }
}
