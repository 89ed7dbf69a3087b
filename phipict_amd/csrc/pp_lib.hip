// Library-wide entry points: error reporting, version, device synchronisation.
#include "pp_common.hpp"

namespace pp {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *get_error() { return g_last_error.c_str(); }
} // namespace pp

extern "C" {
const char *pp_last_error(void) { return pp::get_error(); }
int pp_version(void) { return 1; }
int pp_device_sync(void) {
    PP_HIP(hipDeviceSynchronize());
    return PP_OK;
}
}
