// errors.cpp -- thread-local last-error string of the C ABI.  Messages follow
// the reference's "Class::method || message" convention (e.g.
// SOLVER/src/core/element/material/elastic/3D/Isotropic3D.cpp:34), so the C++
// facade can rethrow them as std::runtime_error unchanged.
#include <string>

#include "axg.h"

namespace axg {
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace axg

extern "C" const char *axg_last_error(void) { return axg::g_last_error.c_str(); }
