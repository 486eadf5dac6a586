// Error state and version of libdsx (host only).
#include <cstring>
#include <string>

#include "dsx_internal.h"

namespace dsx {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DSX_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
    return DSX_OK;
}

}  // namespace dsx

extern "C" int dsx_version(void) { return (0 << 16) | (1 << 8) | 0; }

extern "C" const char *dsx_last_error(void) { return dsx::g_last_error.c_str(); }
