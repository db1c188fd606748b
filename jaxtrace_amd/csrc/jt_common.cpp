// Library-wide error state and version query.
#include "jt_common.h"
#include <stdarg.h>
#include <stdio.h>

namespace jt {
static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace jt

extern "C" const char* jt_last_error(void) { return jt::g_err; }
extern "C" int32_t jt_version(void) { return 1; }
