/*
 * spac_error.cpp — the thread-local error channel of the C ABI (spac_last_error) and the ABI
 * version; host code only.
 */
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/spac_fk.h"
#include "spac_error.h"

namespace spac {

static thread_local std::string g_last_error;

spac_status set_error(spac_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

} // namespace spac

extern "C" {
const char* spac_last_error(void) { return spac::g_last_error.c_str(); }
int32_t spac_abi_version(void) { return SPAC_ABI_VERSION; }
}
