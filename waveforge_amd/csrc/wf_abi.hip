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

extern "C" int wf_status_check(const uint32_t *status, int32_t n_words, hipStream_t stream) {
    if (!status || n_words < 1) {
        wf::set_error("wf_status_check: bad arguments");
        return WF_ERR_ARG;
    }
    uint32_t buf[256];
    for (int32_t off = 0; off < n_words; off += 256) {
        const int32_t n = n_words - off < 256 ? n_words - off : 256;
        hipError_t e = hipMemcpyAsync(buf, status + off, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return wf::check_hip(e, "wf_status_check");
        for (int32_t k = 0; k < n; ++k) {
            const uint32_t v = buf[k];
            if (v == 0u) continue;
            wf::set_error("status word %d = 0x%x:%s%s%s", off + k, v,
                          (v & WF_STATUS_WAVE_END) ? " a source stepped past the end of its waveform table" : "",
                          (v & WF_STATUS_TILE_GROUPS) ? " more than 64 source groups in one tile region" : "",
                          (v & WF_STATUS_TILE_DETS) ? " more than 64 detectors in one tile core" : "");
            return WF_ERR_ARG;
        }
    }
    return WF_OK;
}
