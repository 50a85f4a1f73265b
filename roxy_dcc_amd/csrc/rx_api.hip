// rx_api.hip — library-wide entry points: version, thread-local errors and
// device queries.
#include <string.h>

#include "rx_common.h"

namespace {
thread_local char g_err[1024] = "";
}

namespace rx {

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int device_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

}  // namespace rx

extern "C" int rx_version(void) { return RX_ABI_VERSION; }

extern "C" const char* rx_last_error(void) { return g_err; }
