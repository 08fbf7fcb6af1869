// C-ABI housekeeping: version and thread-local error reporting.
#include "wf_common.hpp"

namespace wf {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return WF_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return WF_ERR_HIP;
}

}  // namespace wf

extern "C" int wf_abi_version(void) { return WF_ABI_VERSION; }

extern "C" const char *wf_last_error(void) { return wf::g_last_error.c_str(); }
