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

// One pixel p of one image (0 <= p < nx_img*ny_img).
template <bool kReg>
__device__ __forceinline__ void das_pixel(const wf_das2d_args &a, int64_t p) {
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

template <bool kReg>
__global__ void __launch_bounds__(256) k_das(wf_das2d_args a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)a.nx_img * a.ny_img) return;
    das_pixel<kReg>(a, p);
}

// Many images in one launch: blockIdx.y picks the image, whose descriptor
// travels in the kernel arguments; blocks past an image's last pixel return
// at once.
constexpr int kBatch = 32;  // descriptors per launch (32 x 96 B of kernel arguments)

struct das_batch {
    wf_das2d_args img[kBatch];
};

__global__ void __launch_bounds__(256) k_das_batch(das_batch b) {
    const wf_das2d_args &a = b.img[blockIdx.y];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)a.nx_img * a.ny_img) return;
    if (a.n_rx <= kRegRx)
        das_pixel<true>(a, p);
    else
        das_pixel<false>(a, p);
}

}  // namespace

static int das_validate(const wf_das2d_args *args, const char *what) {
    if (!args || !args->traces || !args->ant_x || !args->ant_y || !args->image) {
        wf::set_error("%s: null pointer", what);
        return WF_ERR_ARG;
    }
    const wf_das2d_args &a = *args;
    if (a.n_tx < 1 || a.n_rx < 1 || a.n_t < 1 || a.nx_img < 1 || a.ny_img < 1 || a.gate_start < 0) {
        wf::set_error("%s: bad sizes (n_tx=%d n_rx=%d n_t=%d image=%dx%d gate=%d)", what, a.n_tx, a.n_rx, a.n_t,
                      a.nx_img, a.ny_img, a.gate_start);
        return WF_ERR_ARG;
    }
    if (!(a.v_medium > 0.0) || !(a.dt > 0.0)) {
        wf::set_error("%s: v_medium and dt must be > 0", what);
        return WF_ERR_ARG;
    }
    if (((int64_t)a.nx_img * a.ny_img + 255) / 256 > 0x7fffffff) {
        wf::set_error("%s: image too large", what);
        return WF_ERR_ARG;
    }
    return WF_OK;
}

extern "C" int wf_das2d(const wf_das2d_args *args, hipStream_t stream) {
    if (das_validate(args, "wf_das2d") != WF_OK) return WF_ERR_ARG;
    const wf_das2d_args &a = *args;
    const int64_t blocks = ((int64_t)a.nx_img * a.ny_img + 255) / 256;
    if (a.n_rx <= kRegRx)
        hipLaunchKernelGGL(k_das<true>, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(k_das<false>, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return wf::check_launch("wf_das2d");
}

extern "C" int wf_das2d_batch(const wf_das2d_args *args, int32_t n_images, hipStream_t stream) {
    if (!args || n_images < 1) {
        wf::set_error("wf_das2d_batch: bad arguments (args %p, n_images %d)", (const void *)args, n_images);
        return WF_ERR_ARG;
    }
    for (int32_t k = 0; k < n_images; ++k)
        if (das_validate(args + k, "wf_das2d_batch") != WF_OK) return WF_ERR_ARG;
    for (int32_t k0 = 0; k0 < n_images; k0 += kBatch) {
        const int n = n_images - k0 < kBatch ? n_images - k0 : kBatch;
        das_batch b{};
        int64_t blocks = 1;
        for (int k = 0; k < n; ++k) {
            b.img[k] = args[k0 + k];
            const int64_t nb = ((int64_t)b.img[k].nx_img * b.img[k].ny_img + 255) / 256;
            blocks = nb > blocks ? nb : blocks;
        }
        hipLaunchKernelGGL(k_das_batch, dim3((unsigned)blocks, (unsigned)n), dim3(256), 0, stream, b);
        const int rc = wf::check_launch("wf_das2d_batch");
        if (rc != WF_OK) return rc;
    }
    return WF_OK;
}
