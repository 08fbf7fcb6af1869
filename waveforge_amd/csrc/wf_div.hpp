// fp32 division by a fixed divisor (the stepper's and the adjoint's
// divisions by the cell sizes dx, dy), in one of three modes:
//   kDivIeee  IEEE x/d: torch on the CPU divides a float32 tensor by a Python
//             float this way;
//   kDivFast  the exact fast quotient below, equal to IEEE x/d;
//   kDivRcp   x*r with r = fp32(1/d) formed from the double d on the host:
//             torch on a GPU divides a float32 tensor by a Python float this
//             way (its reciprocal-multiply path for a CPU-scalar divisor).
//
// For d = m*2^e and r = RN(1/d), q0 = x*r is corrected by the residual of the
// rescaled problem (x*2^k) - q0*(d*2^k), k = 31 - e, formed with one fma and
// scaled back by r*2^-k: for |x| below the host's span the result equals
// IEEE x/d, which wf_div_verify checks exhaustively on the device for each
// divisor before a caller enables it. Outside the span a wavefront takes the
// IEEE division for the lanes concerned.
#pragma once

#include <hip/hip_runtime.h>

namespace wf {

constexpr int kDivIeee = 0, kDivFast = 1, kDivRcp = 2;

// Divisors the fast quotient takes: positive normal with exponent in [-95, 31].
__host__ __device__ inline bool divider_ok(float d) {
    const unsigned int be = (__builtin_bit_cast(unsigned int, d) >> 23) & 0xffu;
    return d > 0.0f && be >= 32u && be <= 158u;
}

struct Divider {
    float d, r, d2, rk, s, h;
};

// d = m*2^e (normal, -95 <= e <= 31; the host only enables FAST then), r =
// RN(1/d), and the verified range |x| < h with h = bits(span). Every factor is
// a power-of-two rescaling, formed on the exponent bits.
__device__ __forceinline__ Divider make_divider(float d, float r, unsigned int span) {
    const unsigned int db = __float_as_uint(d);
    const unsigned int k = 158u - ((db >> 23) & 0xffu);  // 31 - e
    Divider v;
    v.d = d;
    v.r = r;
    v.d2 = __uint_as_float(db + (k << 23));
    v.rk = __uint_as_float(__float_as_uint(r) - (k << 23));
    v.s = __uint_as_float((k + 127u) << 23);
    v.h = __uint_as_float(span);
    return v;
}

__device__ __forceinline__ float fast_quotient(float x, const Divider &v) {
    const float q0 = x * v.r;
    const float t = __builtin_fmaf(q0, v.d2, -(x * v.s));
    return __builtin_fmaf(-t, v.rk, q0);
}

// x / d. FAST: the scaled corrected quotient for every lane; if any lane of
// the wavefront holds an x outside the verified range, the whole wavefront
// takes the IEEE division and those lanes keep its result. The ballot makes
// the fallback a uniform (scalar) branch.
template <bool FAST>
__device__ __forceinline__ float divide(float x, const Divider &v) {
    if (!FAST) return x / v.d;
    float q = fast_quotient(x, v);
    const bool slow = !(__builtin_fabsf(x) < v.h);
    if (__builtin_expect(__ballot(slow) != 0ull, 0)) {
        const float qs = x / v.d;
        q = slow ? qs : q;
    }
    return q;
}

// The 2*N quotients of one Faraday or Ampere row (N columns per lane, two
// divisions per cell) with one range ballot and one branch between them: the
// stepper's instruction stream is as much scalar as vector, so a branch per
// division would cost as much as its range test. MODE is kDivIeee, kDivFast
// or kDivRcp (v.r is then the host's fp32(1/d) and each quotient one product).
template <int MODE, int N>
__device__ __forceinline__ void divide_row(const float (&x1)[N], const Divider &v1, const float (&x2)[N],
                                           const Divider &v2, float (&q1)[N], float (&q2)[N]) {
    if constexpr (MODE == kDivRcp) {
#pragma unroll
        for (int h = 0; h < N; ++h) {
            q1[h] = x1[h] * v1.r;
            q2[h] = x2[h] * v2.r;
        }
        return;
    }
    if constexpr (MODE == kDivIeee) {
#pragma unroll
        for (int h = 0; h < N; ++h) {
            q1[h] = x1[h] / v1.d;
            q2[h] = x2[h] / v2.d;
        }
        return;
    }
    bool any = false;
#pragma unroll
    for (int h = 0; h < N; ++h) {
        q1[h] = fast_quotient(x1[h], v1);
        q2[h] = fast_quotient(x2[h], v2);
        any = any || !(__builtin_fabsf(x1[h]) < v1.h) || !(__builtin_fabsf(x2[h]) < v2.h);
    }
    if (__builtin_expect(__ballot(any) != 0ull, 0)) {
#pragma unroll
        for (int h = 0; h < N; ++h) {
            const float qa = x1[h] / v1.d, qb = x2[h] / v2.d;
            q1[h] = !(__builtin_fabsf(x1[h]) < v1.h) ? qa : q1[h];
            q2[h] = !(__builtin_fabsf(x2[h]) < v2.h) ? qb : q2[h];
        }
    }
}

}  // namespace wf
