// ep_common.h -- status/error plumbing shared by the host side of the library.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "../../include/eplace.h"

namespace ep {

// Thread-local last-error text (ep_last_error); set by fail().
char *last_error_buffer();

inline ep_status fail(ep_status code, const char *fmt, ...)
{
    char *buf = last_error_buffer();
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace ep
