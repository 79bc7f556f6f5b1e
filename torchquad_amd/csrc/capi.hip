// Library-wide C ABI entry points: error text, version, device count.
#include "tq_host.hpp"

namespace tq {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

}  // namespace tq

extern "C" {

const char* tq_last_error(void) { return tq::g_last_error; }

int tq_version(void) { return 100; }

int tq_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
