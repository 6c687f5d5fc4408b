/*
 * spac_error.h — the thread-local error channel behind spac_last_error().
 */
#ifndef SPAC_ERROR_H
#define SPAC_ERROR_H

#include "../../include/spac_fk.h"

namespace spac {
spac_status set_error(spac_status code, const char* fmt, ...);
}

#endif
