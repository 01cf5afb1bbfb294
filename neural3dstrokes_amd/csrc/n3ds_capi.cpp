// Error state and version query for the C-ABI.
#include <stdarg.h>

#include "n3ds_capi.h"

namespace n3ds {

static thread_local char g_error[1024] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

}  // namespace n3ds

N3DS_EXPORT const char* n3ds_last_error(void) { return n3ds::g_error; }

N3DS_EXPORT int n3ds_version(void) { return 1; }
