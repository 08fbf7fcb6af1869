// Shared host/device helpers for libwaveforge_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdarg>
#include <string>

#include "waveforge_hip.h"

namespace wf {

// Thread-local last error message returned by wf_last_error().
void set_error(const char *fmt, ...);

// Record a HIP error (if any) from the most recent launch or call.
int check_hip(hipError_t e, const char *what);
inline int check_launch(const char *what) { return check_hip(hipGetLastError(), what); }

// Reinterpret a non-negative float's bit pattern for atomicMax-based maxima:
// |x| bit patterns order like their values, and NaN patterns sort above +inf,
// so a NaN anywhere wins the maximum as torch's max() does.
__device__ __forceinline__ unsigned int abs_bits(float x) {
    return __float_as_uint(x) & 0x7fffffffu;
}

}  // namespace wf
