// Reference-exact per-step path: three launches per leapfrog step, in place.
//
//   k_pre  (1 workgroup): Mur snapshot of the rows {0,1,nx-2,nx-1} and columns
//          {0,1,ny-2,ny-1} of Hz (the only parts of MurABC.Hz_prev that
//          MurABC.apply reads, src/core/boundaries.py:259-274), then the soft
//          sources in collection order (src/core/sources.py:1060-1071).
//   k_h    Faraday on interior cells (src/core/fdtd2d.py:232-235). Each Mur
//          edge cell is written by the interior thread whose Faraday value it
//          needs, so the launch is race-free in place: (1,j) writes (0,j),
//          (nx-2,j) writes (nx-1,j), (i,1) writes (i,0), (i,ny-2) writes
//          (i,ny-1), and the four corners go to (1,1), (1,ny-2), (nx-2,1),
//          (nx-2,ny-2). Faraday values on row 0 / column 0 are never needed:
//          Mur overwrites them.
//   k_e    Ampere on every cell (src/core/fdtd2d.py:244-257) plus the Hz
//          detector gather the reference's callers do after step()
//          (examples/brain_mimo_imaging.py:196-198).
//
// All arithmetic is fp32 with one rounding per operation, in the reference's
// operation order; the library is compiled with -ffp-contract=off so no
// multiply-add is fused. Division is IEEE division by fp32(dx), as torch's
// CPU `tensor / python_float` is, or, for a run flagged WF_RUN2D_RECIPROCAL,
// the product by fp32(1/dx) that torch on a GPU forms for the same
// expression.
#include "wf_common.hpp"

#include <mutex>
#include <type_traits>

namespace {

constexpr int kBX = 64;
constexpr int kBY = 4;

// The cell bodies are templates over the run descriptor, so the fp32 path
// (wf_run2d) and the fp64 check path (wf_run2d_f64) share one statement of
// the step; T is the descriptor's field type.
template <class R>
using real_t = typename std::remove_cv<typename std::remove_pointer<decltype(R::ex)>::type>::type;

template <class R>
__device__ __forceinline__ real_t<R> prow(const R &r, int s, int j) {
    return r.pstrip[s * r.ny + j];  // s: 0 -> row 0, 1 -> row 1, 2 -> row nx-2, 3 -> row nx-1
}
template <class R>
__device__ __forceinline__ real_t<R> pcol(const R &r, int s, int i) {
    return r.pstrip[4 * r.ny + s * r.nx + i];  // s: 0 -> col 0, 1 -> col 1, 2 -> col ny-2, 3 -> col ny-1
}

// Snapshot strips then soft sources of step n; one workgroup per run.
template <class R>
__device__ __forceinline__ void pre_body(const R &r, int n) {
    using T = real_t<R>;
    const int nx = r.nx, ny = r.ny;
    const int tid = threadIdx.x;
    for (int t = tid; t < 4 * ny; t += blockDim.x) {
        const int s = t / ny, j = t - s * ny;
        const int i = s < 2 ? s : nx - 4 + s;
        r.pstrip[t] = r.hz[(size_t)i * ny + j];
    }
    for (int t = tid; t < 4 * nx; t += blockDim.x) {
        const int s = t / nx, i = t - s * nx;
        const int j = s < 2 ? s : ny - 4 + s;
        r.pstrip[4 * ny + t] = r.hz[(size_t)i * ny + j];
    }
    __syncthreads();
    for (int g = tid; g < r.n_grp; g += blockDim.x) {
        const int a = r.grp_start[g], b = r.grp_start[g + 1];
        const int comp = r.inj[a].comp, cell = r.inj[a].cell;
        T *f = comp == WF_COMP_EX ? r.ex : (comp == WF_COMP_EY ? r.ey : r.hz);
        T v = f[cell];
        for (int c = a; c < b; ++c) {
            if ((unsigned)n >= (unsigned)r.inj[c].wave_len) {  // past the table: report, inject nothing
                if (r.status) atomicOr(r.status, WF_STATUS_WAVE_END);
                continue;
            }
            const T w = r.inj[c].wave[n] * r.inj[c].amp;  // values = w_now * amp
            v = v + w;                                    // target[idx] += values
        }
        f[cell] = v;
    }
}

// x / dx and x / dy: IEEE division, or the product by the run's reciprocal
// when it is flagged WF_RUN2D_RECIPROCAL (fp32 runs only).
template <class R>
__device__ __forceinline__ real_t<R> over_dx(const R &r, real_t<R> x) {
    if constexpr (std::is_same<R, wf_run2d>::value)
        if (r.flags & WF_RUN2D_RECIPROCAL) return x * r.rdx;
    return x / r.dx;
}
template <class R>
__device__ __forceinline__ real_t<R> over_dy(const R &r, real_t<R> x) {
    if constexpr (std::is_same<R, wf_run2d>::value)
        if (r.flags & WF_RUN2D_RECIPROCAL) return x * r.rdy;
    return x / r.dy;
}

// Faraday at interior cell (i, j) plus the Mur edge cells it owns.
template <class R>
__device__ __forceinline__ void h_body(const R &r, int i, int j) {
    using T = real_t<R>;
    const int nx = r.nx, ny = r.ny;
    const size_t c = (size_t)i * ny + j;
    const T h1 = r.hz[c];
    const T a = over_dy(r, r.ex[c + 1] - r.ex[c]);
    const T b = over_dx(r, r.ey[c + ny] - r.ey[c]);
    const T F = h1 + r.dh * (a - b);
    r.hz[c] = F;
    const T cx = r.cmx, cy = r.cmy;
    if (i == 1) {  // left edge (0, j), then the corners that use it
        const T L = prow(r, 1, j) + cx * (F - prow(r, 0, j));
        r.hz[j] = L;
        if (j == 1) r.hz[0] = prow(r, 0, 1) + cy * (L - prow(r, 0, 0));
        if (j == ny - 2) r.hz[ny - 1] = prow(r, 0, ny - 2) + cy * (L - prow(r, 0, ny - 1));
    }
    if (i == nx - 2) {  // right edge (nx-1, j)
        const T R_ = prow(r, 2, j) + cx * (F - prow(r, 3, j));
        r.hz[(size_t)(nx - 1) * ny + j] = R_;
        if (j == 1) r.hz[(size_t)(nx - 1) * ny] = prow(r, 3, 1) + cy * (R_ - prow(r, 3, 0));
        if (j == ny - 2)
            r.hz[(size_t)(nx - 1) * ny + ny - 1] = prow(r, 3, ny - 2) + cy * (R_ - prow(r, 3, ny - 1));
    }
    if (j == 1) r.hz[(size_t)i * ny] = pcol(r, 1, i) + cy * (F - pcol(r, 0, i));
    if (j == ny - 2) r.hz[(size_t)i * ny + ny - 1] = pcol(r, 2, i) + cy * (F - pcol(r, 3, i));
}

// Detector gather of step n (Hz after the step); threads of one workgroup.
template <class R>
__device__ __forceinline__ void record_body(const R &r, int n) {
    const int t = n - r.trace_n0;
    if (t >= 0 && t < r.trace_len) {
        for (int d = threadIdx.y * kBX + threadIdx.x; d < r.n_det; d += kBX * kBY)
            r.traces[d * r.trace_stride + t] = r.hz[r.det_cell[d]];
    }
}

// Ampere at cell (i, j).
template <bool kMat, class R>
__device__ __forceinline__ void e_body(const R &r, int i, int j) {
    using T = real_t<R>;
    const int ny = r.ny;
    const size_t c = (size_t)i * ny + j;
    const T h = r.hz[c];
    if (j >= 1) {
        const T g = over_dy(r, h - r.hz[c - 1]);
        if (kMat) {
            const T t1 = r.ca[c] * r.ex[c];
            const T t2 = r.cb[c] * g;
            r.ex[c] = t1 + t2;
        } else {
            const T t2 = r.de * g;
            r.ex[c] = r.ex[c] + t2;
        }
    }
    if (i >= 1) {
        const T g = over_dx(r, h - r.hz[c - ny]);
        if (kMat) {
            const T t1 = r.ca[c] * r.ey[c];
            const T t2 = r.cb[c] * g;
            r.ey[c] = t1 - t2;
        } else {
            const T t2 = r.de * g;
            r.ey[c] = r.ey[c] - t2;
        }
    }
}

__global__ void __launch_bounds__(1024) k_pre(wf_run2d r, int n) { pre_body(r, n); }

__global__ void __launch_bounds__(kBX * kBY) k_h(wf_run2d r) {
    const int j = 1 + blockIdx.x * kBX + threadIdx.x;
    const int i = 1 + blockIdx.y * kBY + threadIdx.y;
    if (i > r.nx - 2 || j > r.ny - 2) return;
    h_body(r, i, j);
}

template <bool kMat>
__global__ void __launch_bounds__(kBX * kBY) k_e(wf_run2d r, int n) {
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    if (blockIdx.x == 0 && blockIdx.y == 0) record_body(r, n);
    if (i >= r.nx || j >= r.ny) return;
    e_body<kMat>(r, i, j);
}

// fp64 check path: one launch per stage covers every run (blockIdx.z or
// blockIdx.x is the run); descriptors are read from the device table.
__global__ void __launch_bounds__(1024) k_pre64(const wf_run2d_f64 *runs, int n_off) {
    const wf_run2d_f64 &r = runs[blockIdx.x];
    pre_body(r, r.step + n_off);
}

__global__ void __launch_bounds__(kBX * kBY) k_h64(const wf_run2d_f64 *runs) {
    const wf_run2d_f64 &r = runs[blockIdx.z];
    const int j = 1 + blockIdx.x * kBX + threadIdx.x;
    const int i = 1 + blockIdx.y * kBY + threadIdx.y;
    if (i > r.nx - 2 || j > r.ny - 2) return;
    h_body(r, i, j);
}

__global__ void __launch_bounds__(kBX * kBY) k_e64(const wf_run2d_f64 *runs, int n_off) {
    const wf_run2d_f64 &r = runs[blockIdx.z];
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    if (blockIdx.x == 0 && blockIdx.y == 0) record_body(r, r.step + n_off);
    if (i >= r.nx || j >= r.ny) return;
    if (r.ca)
        e_body<true>(r, i, j);
    else
        e_body<false>(r, i, j);
}

// |x| as an unsigned bit pattern: orders like |x|, and NaN sorts above inf,
// so a NaN anywhere wins the maximum as torch's max() does.
__device__ __forceinline__ unsigned int abs_key(float x) { return wf::abs_bits(x); }
__device__ __forceinline__ unsigned long long abs_key(double x) {
    return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
}
__device__ __forceinline__ float from_key(unsigned int k) { return __uint_as_float(k); }
__device__ __forceinline__ double from_key(unsigned long long k) { return __longlong_as_double((long long)k); }

template <class T, class K>
__global__ void __launch_bounds__(256) k_maxabs_partial(const T *__restrict__ ex, const T *__restrict__ ey,
                                                        const T *__restrict__ hz, int64_t n,
                                                        K *__restrict__ partial) {
    K m[3] = {0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        m[0] = max(m[0], abs_key(ex[k]));
        m[1] = max(m[1], abs_key(ey[k]));
        m[2] = max(m[2], abs_key(hz[k]));
    }
    __shared__ K sm[3][256];
    for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = max(sm[c][threadIdx.x], sm[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

template <class T, class K>
__global__ void __launch_bounds__(256) k_maxabs_final(const K *__restrict__ partial, int np, T *__restrict__ out3) {
    __shared__ K sm[3][256];
    K m[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < np; b += blockDim.x)
        for (int c = 0; c < 3; ++c) m[c] = max(m[c], partial[3 * b + c]);
    for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = max(sm[c][threadIdx.x], sm[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) out3[threadIdx.x] = from_key(sm[threadIdx.x][0]);
}

int validate_run(const wf_run2d &r) {
    if (!r.ex || !r.ey || !r.hz) { wf::set_error("run: null field pointer"); return WF_ERR_ARG; }
    if (r.nx < 4 || r.ny < 4) { wf::set_error("run: grid must be at least 4x4, got %dx%d", r.nx, r.ny); return WF_ERR_ARG; }
    if ((int64_t)r.nx * r.ny > INT32_MAX) {  // cells are int32 in the source and detector tables
        wf::set_error("run: %dx%d cells exceed the 2^31-cell index range of the source and detector tables", r.nx, r.ny);
        return WF_ERR_UNSUPPORTED;
    }
    if ((r.ca == nullptr) != (r.cb == nullptr)) { wf::set_error("run: ca and cb must both be set or both NULL"); return WF_ERR_ARG; }
    if (r.n_grp < 0 || r.n_det < 0) { wf::set_error("run: negative table size"); return WF_ERR_ARG; }
    if (r.n_grp > 0 && (!r.inj || !r.grp_start)) { wf::set_error("run: source tables missing"); return WF_ERR_ARG; }
    if (r.n_det > 0 && (!r.det_cell || !r.traces)) { wf::set_error("run: detector tables missing"); return WF_ERR_ARG; }
    return WF_OK;
}

// The launches of n_steps steps of every run (validated by the caller). With
// `strips` set, every run takes its Mur snapshot strips there instead of its
// own pstrip: the runs' steps are stream-ordered (run b's k_pre, k_h, k_e
// before run b+1's), so one strip buffer serves them all in turn.
int launch_steps(const wf_run2d *runs, int n_runs, int n_steps, float *strips, hipStream_t stream, const char *what) {
    for (int s = 0; s < n_steps; ++s) {
        for (int b = 0; b < n_runs; ++b) {
            wf_run2d r = runs[b];
            if (strips) r.pstrip = strips;
            const int n = r.step + s;
            const int pre_threads = 1024;
            hipLaunchKernelGGL(k_pre, dim3(1), dim3(pre_threads), 0, stream, r, n);
            dim3 gh((r.ny - 2 + kBX - 1) / kBX, (r.nx - 2 + kBY - 1) / kBY);
            hipLaunchKernelGGL(k_h, gh, dim3(kBX, kBY), 0, stream, r);
            dim3 ge((r.ny + kBX - 1) / kBX, (r.nx + kBY - 1) / kBY);
            if (r.ca)
                hipLaunchKernelGGL(k_e<true>, ge, dim3(kBX, kBY), 0, stream, r, n);
            else
                hipLaunchKernelGGL(k_e<false>, ge, dim3(kBX, kBY), 0, stream, r, n);
        }
        const int rc = wf::check_launch(what);
        if (rc) return rc;
    }
    return WF_OK;
}

// Scratch of the allocation-free step: the Mur snapshot strips of the run being
// stepped, 4*(nx+ny) floats, in a file-scope device array. Each GPU holds its
// own copy until the process ends; in host code the array's name is not a
// device address, so the address is taken once per device with
// hipGetSymbolAddress and only that address is used.
constexpr int64_t kNoAllocStripFloats = int64_t(1) << 18;  // nx + ny <= 65536
__device__ __align__(256) float scratch[kNoAllocStripFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(float **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return wf::check_hip(e, "wf_fdtd2d_step_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        wf::set_error("wf_fdtd2d_step_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return WF_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return wf::check_hip(e, "wf_fdtd2d_step_noalloc: hipGetSymbolAddress");
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return WF_OK;
}

}  // namespace

extern "C" int wf_fdtd2d_step(const wf_run2d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream) {
    if (!runs || n_runs < 1 || n_steps < 0) { wf::set_error("wf_fdtd2d_step: bad arguments"); return WF_ERR_ARG; }
    for (int b = 0; b < n_runs; ++b) {
        const int rc = validate_run(runs[b]);
        if (rc) return rc;
        if (!runs[b].pstrip) { wf::set_error("wf_fdtd2d_step: pstrip workspace is NULL"); return WF_ERR_ARG; }
    }
    return launch_steps(runs, n_runs, n_steps, nullptr, stream, "wf_fdtd2d_step");
}

extern "C" int wf_fdtd2d_step_noalloc(const wf_run2d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream) {
    if (!runs || n_runs < 1 || n_steps < 0) {
        wf::set_error("wf_fdtd2d_step_noalloc: bad arguments");
        return WF_ERR_ARG;
    }
    for (int b = 0; b < n_runs; ++b) {
        const int rc = validate_run(runs[b]);
        if (rc) return rc;
        if (4 * ((int64_t)runs[b].nx + runs[b].ny) > kNoAllocStripFloats) {
            wf::set_error("wf_fdtd2d_step_noalloc: run %d (%dx%d) needs %lld strip floats, above the %lld of the "
                          "library's scratch (nx + ny <= %lld)", b, runs[b].nx, runs[b].ny,
                          (long long)(4 * ((int64_t)runs[b].nx + runs[b].ny)), (long long)kNoAllocStripFloats,
                          (long long)(kNoAllocStripFloats / 4));
            return WF_ERR_ARG;
        }
    }
    float *strips = nullptr;
    const int rc = scratch_for_current_device(&strips);
    if (rc) return rc;
    return launch_steps(runs, n_runs, n_steps, strips, stream, "wf_fdtd2d_step_noalloc");
}

extern "C" int wf_fdtd2d_maxabs(const float *ex, const float *ey, const float *hz, int64_t n, float *partial,
                                int32_t n_partial, float *out3, hipStream_t stream) {
    if (!ex || !ey || !hz || !partial || !out3 || n <= 0 || n_partial < 1 || n_partial > 4096) {
        wf::set_error("wf_fdtd2d_maxabs: bad arguments");
        return WF_ERR_ARG;
    }
    unsigned int *p = reinterpret_cast<unsigned int *>(partial);
    hipLaunchKernelGGL((k_maxabs_partial<float, unsigned int>), dim3(n_partial), dim3(256), 0, stream, ex, ey, hz,
                       n, p);
    hipLaunchKernelGGL((k_maxabs_final<float, unsigned int>), dim3(1), dim3(256), 0, stream, p, n_partial, out3);
    return wf::check_launch("wf_fdtd2d_maxabs");
}

extern "C" int wf_fdtd2d_maxabs_f64(const double *ex, const double *ey, const double *hz, int64_t n,
                                    double *partial, int32_t n_partial, double *out3, hipStream_t stream) {
    if (!ex || !ey || !hz || !partial || !out3 || n <= 0 || n_partial < 1 || n_partial > 4096) {
        wf::set_error("wf_fdtd2d_maxabs_f64: bad arguments");
        return WF_ERR_ARG;
    }
    unsigned long long *p = reinterpret_cast<unsigned long long *>(partial);
    hipLaunchKernelGGL((k_maxabs_partial<double, unsigned long long>), dim3(n_partial), dim3(256), 0, stream, ex,
                       ey, hz, n, p);
    hipLaunchKernelGGL((k_maxabs_final<double, unsigned long long>), dim3(1), dim3(256), 0, stream, p, n_partial,
                       out3);
    return wf::check_launch("wf_fdtd2d_maxabs_f64");
}

extern "C" int wf_fdtd2d_step_f64(const wf_f64_args *args, hipStream_t stream) {
    if (!args || !args->runs || args->n_runs < 1 || args->n_runs > 65535 || args->n_steps < 0 ||
        args->max_nx < 4 || args->max_ny < 4) {
        wf::set_error("wf_fdtd2d_step_f64: bad arguments");
        return WF_ERR_ARG;
    }
    const wf_f64_args &a = *args;
    dim3 gh((a.max_ny - 2 + kBX - 1) / kBX, (a.max_nx - 2 + kBY - 1) / kBY, a.n_runs);
    dim3 ge((a.max_ny + kBX - 1) / kBX, (a.max_nx + kBY - 1) / kBY, a.n_runs);
    for (int s = 0; s < a.n_steps; ++s) {
        const int off = a.step_offset + s;
        hipLaunchKernelGGL(k_pre64, dim3(a.n_runs), dim3(1024), 0, stream, a.runs, off);
        hipLaunchKernelGGL(k_h64, gh, dim3(kBX, kBY), 0, stream, a.runs);
        hipLaunchKernelGGL(k_e64, ge, dim3(kBX, kBY), 0, stream, a.runs, off);
        if ((s & 255) == 255 || s + 1 == a.n_steps) {
            const int rc = wf::check_launch("wf_fdtd2d_step_f64");
            if (rc) return rc;
        }
    }
    return WF_OK;
}
