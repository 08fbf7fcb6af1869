// Delay-and-sum backprojection (examples/brain_mimo_imaging.py:211-266 and
// variants), one thread per pixel.
//
// Every value is formed exactly as NumPy forms it in the reference, so the
// image matches it bit for bit:
//   px = u * pitch_x                                       (fp64)
//   d  = sqrt((px - ax)^2 + (py - ay)^2)                   (fp64, x*x squares)
//   n  = int32(rint(((d_tx + d_rx) / v) / dt))             (half-even rounding)
//   img += float64(S[tx, rx, clip(n)])  in tx-major, rx-minor order
//   out = img * img
// The library is built with -ffp-contract=off, so no product is fused into a
// sum; fp64 sqrt and division are the correctly rounded device routines.
//
// Work per image: nx_img*ny_img*n_tx*n_rx gathers. The reference recomputes
// the pixel's distance to each receive antenna for every pair; the same fp64
// operations on the same inputs give the same value, so for arrays of up to
// kRegRx elements each thread forms the n_rx distances once, in registers
// (n_tx + n_rx square roots per pixel instead of n_tx*(n_rx + 1)). The trace
// set (n_tx*n_rx*n_t*4 bytes, 2 MB for 16x16x2000) stays in L2, and the 64
// neighbouring pixels of a wavefront read nearby samples.
#include "wf_common.hpp"

namespace {

constexpr int kRegRx = 32;  // receive elements whose distances a thread keeps in registers

__device__ __forceinline__ double distance(double px, double py, double ax, double ay) {
    const double ex = px - ax;
    const double ey = py - ay;
    return sqrt(ex * ex + ey * ey);
}

// One pair's sample: n = int32(rint(((d_tx + d_rx) / v) / dt)), clipped or
// zeroed outside [0, n_t), zero before the gate.
__device__ __forceinline__ float sample(const wf_das2d_args &a, const float *row, double dtx, double drx) {
    const int nt = a.n_t;
    const double tau = (dtx + drx) / a.v_medium;
    const int n = (int)rint(tau / a.dt);
    if (a.clip) {
        const int nc = n < 0 ? 0 : (n > nt - 1 ? nt - 1 : n);
        return nc < a.gate_start ? 0.0f : row[nc];
    }
    return (n < 0 || n >= nt || n < a.gate_start) ? 0.0f : row[n];
}

template <bool kReg>
__global__ void __launch_bounds__(256) k_das(wf_das2d_args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npx = (int64_t)a.nx_img * a.ny_img;
    if (p >= npx) return;
    const int u = (int)(p / a.ny_img);
    const int v = (int)(p - (int64_t)u * a.ny_img);
    const double px = (double)u * a.pitch_x;
    const double py = (double)v * a.pitch_y;
    const int nt = a.n_t;
    double drx_r[kReg ? kRegRx : 1];
    if (kReg) {
#pragma unroll
        for (int rx = 0; rx < kRegRx; ++rx)
            drx_r[rx] = rx < a.n_rx ? distance(px, py, a.ant_x[rx], a.ant_y[rx]) : 0.0;
    }
    double img = 0.0;
    for (int tx = 0; tx < a.n_tx; ++tx) {
        const double dtx = distance(px, py, a.ant_x[tx], a.ant_y[tx]);
        const float *row = a.traces + (int64_t)tx * a.n_rx * nt;
        if (kReg) {
#pragma unroll
            for (int rx = 0; rx < kRegRx; ++rx)
                if (rx < a.n_rx) img = img + (double)sample(a, row + (int64_t)rx * nt, dtx, drx_r[rx]);
        } else {
            for (int rx = 0; rx < a.n_rx; ++rx) {
                const double drx = distance(px, py, a.ant_x[rx], a.ant_y[rx]);
                img = img + (double)sample(a, row + (int64_t)rx * nt, dtx, drx);
            }
        }
    }
    a.image[p] = img * img;
}

}  // namespace

extern "C" int wf_das2d(const wf_das2d_args *args, hipStream_t stream) {
    if (!args || !args->traces || !args->ant_x || !args->ant_y || !args->image) {
        wf::set_error("wf_das2d: null pointer");
        return WF_ERR_ARG;
    }
    const wf_das2d_args &a = *args;
    if (a.n_tx < 1 || a.n_rx < 1 || a.n_t < 1 || a.nx_img < 1 || a.ny_img < 1 || a.gate_start < 0) {
        wf::set_error("wf_das2d: bad sizes (n_tx=%d n_rx=%d n_t=%d image=%dx%d gate=%d)", a.n_tx, a.n_rx, a.n_t,
                      a.nx_img, a.ny_img, a.gate_start);
        return WF_ERR_ARG;
    }
    if (!(a.v_medium > 0.0) || !(a.dt > 0.0)) {
        wf::set_error("wf_das2d: v_medium and dt must be > 0");
        return WF_ERR_ARG;
    }
    const int64_t npx = (int64_t)a.nx_img * a.ny_img;
    const int64_t blocks = (npx + 255) / 256;
    if (blocks > 0x7fffffff) {
        wf::set_error("wf_das2d: image too large");
        return WF_ERR_ARG;
    }
    if (a.n_rx <= kRegRx)
        hipLaunchKernelGGL(k_das<true>, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(k_das<false>, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return wf::check_launch("wf_das2d");
}
