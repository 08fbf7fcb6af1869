// Envelope delay-and-sum image with layered-medium travel times: the
// dataset generator's DAS (AntennaRing.compute_das_image,
// src/datasets/brain/antenna.py:262-333). Three stages, each one launch:
//
//   k_env     |analytic signal| of every trace. The analytic signal of an
//             n-sample record (SciPy's FFT definition: spectrum times
//             1, 2, ..., 2, [1], 0, ...) has real part x and imaginary part
//             the circular convolution of x with a fixed kernel k (host
//             supplied, fp64). One workgroup computes 1024 outputs of one
//             trace; x and the k window it needs are staged through LDS in
//             1024-sample chunks; accumulation in fp64.
//   k_tt      one-way travel time in samples for every (antenna, pixel), fp64
//             in the reference's expression order (ray-circle chords through
//             concentric layers, or straight free-space), stored as fp32.
//   k_img     per pixel, the fp64 sum over tx-major, rx-minor pairs of
//             env[tx, rx][rint(tt[tx] + tt[rx])] (fp32 add, half-even rint;
//             indices outside [0, n_t) add nothing), then img^2 as fp32.
//
// The library is built with -ffp-contract=off: every fp64 operation of k_tt
// rounds once, in the reference's order, so the table equals the one the
// reference's Python loops build.
#include "wf_common.hpp"

namespace {

constexpr int kEnvThreads = 256;
constexpr int kEnvPer = 4;                      // outputs per thread
constexpr int kEnvOut = kEnvThreads * kEnvPer;  // outputs per workgroup
constexpr int kChunk = 1024;                    // input samples staged per pass

__global__ void __launch_bounds__(kEnvThreads) k_env(const float *__restrict__ traces,
                                                     const double *__restrict__ kern,
                                                     float *__restrict__ env, int n_t) {
    __shared__ double xs[kChunk];
    __shared__ double kw[kEnvOut + kChunk];
    const int tid = threadIdx.x;
    const int64_t trace = blockIdx.y;
    const float *x = traces + trace * n_t;
    const int n0 = blockIdx.x * kEnvOut;
    double acc[kEnvPer];
#pragma unroll
    for (int r = 0; r < kEnvPer; ++r) acc[r] = 0.0;
    for (int m0 = 0; m0 < n_t; m0 += kChunk) {
        __syncthreads();
        for (int t = tid; t < kChunk; t += kEnvThreads) xs[t] = (m0 + t < n_t) ? (double)x[m0 + t] : 0.0;
        // kw[t] = k[(n0 - m0 - (kChunk - 1) + t) mod n_t]
        const int base = n0 - m0 - (kChunk - 1);
        for (int t = tid; t < kEnvOut + kChunk; t += kEnvThreads) {
            int d = (base + t) % n_t;
            if (d < 0) d += n_t;
            kw[t] = kern[d];
        }
        __syncthreads();
        const int lim = min(kChunk, n_t - m0);
        for (int mm = 0; mm < lim; ++mm) {
            const double xv = xs[mm];
#pragma unroll
            for (int r = 0; r < kEnvPer; ++r) {
                // n - m = (n0 + tid + r*256) - (m0 + mm)  ->  kw index tid + r*256 - mm + kChunk - 1
                acc[r] += xv * kw[tid + r * kEnvThreads - mm + kChunk - 1];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kEnvPer; ++r) {
        const int n = n0 + tid + r * kEnvThreads;
        if (n < n_t) {
            const double re = (double)x[n];
            env[trace * n_t + n] = (float)sqrt(re * re + acc[r] * acc[r]);
        }
    }
}

constexpr double kC0Ray = 3e8;

__device__ __forceinline__ double pymax0(double v) { return v > 0.0 ? v : 0.0; }        // max(0.0, v)
__device__ __forceinline__ double pymin(double a, double b) { return b < a ? b : a; }    // min(a, b)

__device__ double chord(double ax, double ay, double bx, double by, double cx, double cy, double r) {
    const double dx_ = bx - ax;
    const double dy_ = by - ay;
    const double seg_len = sqrt(dx_ * dx_ + dy_ * dy_) + 1e-30;
    const double ux = dx_ / seg_len, uy = dy_ / seg_len;
    const double ox = ax - cx, oy = ay - cy;
    const double b_c = ox * ux + oy * uy;
    const double c_c = ox * ox + oy * oy - r * r;
    const double disc = b_c * b_c - c_c;
    if (disc <= 0.0) return 0.0;
    const double sq = sqrt(disc);
    const double t0 = pymax0(pymin(seg_len, -b_c - sq));
    const double t1 = pymax0(pymin(seg_len, -b_c + sq));
    return pymax0(t1 - t0);
}

__global__ void __launch_bounds__(256) k_tt(wf_das_env_args a) {
    const int64_t npx = (int64_t)a.img_w * a.img_h;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= npx * a.n_ant) return;
    const int ant = (int)(gid / npx);
    const int64_t p = gid - (int64_t)ant * npx;
    const int iy = (int)(p / a.img_w), ix = (int)(p - (int64_t)iy * a.img_w);
    const double ax = a.ant_x[ant], ay = a.ant_y[ant];
    const double bx = a.px[ix], by = a.py[iy];
    double tsamp;
    if (a.n_layers > 0) {
        const double total_d = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay)) + 1e-30;
        double prev = 0.0, t = 0.0;
        for (int l = 0; l < a.n_layers; ++l) {
            const double co = chord(ax, ay, bx, by, a.cx, a.cy, a.layer_r[l]);
            const double ci = (l + 1 < a.n_layers) ? chord(ax, ay, bx, by, a.cx, a.cy, a.layer_r[l + 1]) : 0.0;
            const double sc = pymax0(co - ci);
            t += (sc * 1e-3) / (kC0Ray / sqrt(a.layer_eps[l]));
            prev += sc;
        }
        t += (pymax0(total_d - prev) * 1e-3) / kC0Ray;
        tsamp = t / a.dt;
    } else {
        const double d = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
        tsamp = (d * 1e-3) / kC0Ray / a.dt;
    }
    a.tt[gid] = (float)tsamp;
}

__global__ void __launch_bounds__(256) k_img(wf_das_env_args a) {
    const int64_t npx = (int64_t)a.img_w * a.img_h;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    double acc = 0.0;
    for (int tx = 0; tx < a.n_ant; ++tx) {
        const float ttx = a.tt[(int64_t)tx * npx + p];
        const float *e_tx = a.env + (int64_t)tx * a.n_ant * a.n_t;
        for (int rx = 0; rx < a.n_ant; ++rx) {
            const float s = ttx + a.tt[(int64_t)rx * npx + p];
            const float rs = rintf(s);
            if (rs >= 0.0f && rs < (float)a.n_t) acc += (double)e_tx[(int64_t)rx * a.n_t + (int)rs];
        }
    }
    a.image[p] = (float)(acc * acc);
}

}  // namespace

extern "C" int wf_das_envelope(const wf_das_env_args *args, hipStream_t stream) {
    if (!args) {
        wf::set_error("wf_das_envelope: null args");
        return WF_ERR_ARG;
    }
    const wf_das_env_args &a = *args;
    if (a.n_ant < 1 || a.n_t < 1 || a.img_w < 1 || a.img_h < 1 || a.n_layers < 0 || a.n_ant > 65535 ||
        (int64_t)a.n_ant * a.n_ant > 65535 || a.stages < 1 || a.stages > 7 || !(a.dt > 0.0)) {
        wf::set_error("wf_das_envelope: bad sizes");
        return WF_ERR_ARG;
    }
    if (((a.stages & 1) && (!a.traces || !a.hilbert_k || !a.env)) ||
        ((a.stages & 2) && (!a.tt || !a.ant_x || !a.ant_y || !a.px || !a.py ||
                            (a.n_layers > 0 && (!a.layer_r || !a.layer_eps)))) ||
        ((a.stages & 4) && (!a.env || !a.tt || !a.image))) {
        wf::set_error("wf_das_envelope: null pointer for a requested stage");
        return WF_ERR_ARG;
    }
    if (a.stages & 1) {
        dim3 g((a.n_t + kEnvOut - 1) / kEnvOut, a.n_ant * a.n_ant);
        hipLaunchKernelGGL(k_env, g, dim3(kEnvThreads), 0, stream, a.traces, a.hilbert_k, a.env, a.n_t);
    }
    const int64_t npx = (int64_t)a.img_w * a.img_h;
    if (a.stages & 2) {
        const int64_t n = npx * a.n_ant;
        hipLaunchKernelGGL(k_tt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a);
    }
    if (a.stages & 4) hipLaunchKernelGGL(k_img, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, stream, a);
    return wf::check_launch("wf_das_envelope");
}
