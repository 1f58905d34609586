// The error channel of the library (s2h::set_error, s2h::fail, s2h_last_error) for the
// host units built on their own, outside the HIP library, by the sanitizer
// builds of this folder.
#include <string>

#include "../../include/s2hip.h"

namespace s2h {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
s2h_status fail(s2h_status code, const std::string &msg) {
    g_error = msg;
    return code;
}
}  // namespace s2h

extern "C" const char *s2h_last_error(void) { return s2h::g_error.c_str(); }
