// C-ABI entry points of the temporally blocked stepper (kernels in
// wf_tb_kernel.hpp, instantiated by wf_tb_{mat,free}_{fast,ieee}.hip) and the
// exhaustive division check that licenses its fast quotient.
#include "wf_tb_kernel.hpp"

namespace {

// (tile_i, columns per lane) of every compiled shape
constexpr int kTileShapes[][2] = {{32, 1}, {16, 1}, {48, 1}, {64, 1}, {16, 2}, {32, 2}, {48, 2}, {8, 1}};

// Exhaustive check of the fast quotient (wf_tb::divide<true>) against IEEE
// division for every x with lo <= |x bits| < lo + span (both signs). Each
// thread covers a strided slice and counts mismatches.
__global__ void __launch_bounds__(256) k_div_verify(float d, float r, unsigned int lo, unsigned int span,
                                                     unsigned int *bad) {
    const unsigned int nthreads = gridDim.x * blockDim.x;
    const unsigned int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const wf_tb::Divider v = wf_tb::make_divider(d, r, 0x7f800000u);
    unsigned int local = 0;
    for (unsigned int k = tid; k < span; k += nthreads) {
        const float x = __uint_as_float(lo + k);
        const float q = x / d;
        const float qn = (-x) / d;
        if (__float_as_uint(wf_tb::divide<true>(x, v)) != __float_as_uint(q)) ++local;
        if (__float_as_uint(wf_tb::divide<true>(-x, v)) != __float_as_uint(qn)) ++local;
    }
    if (local) atomicAdd(bad, local);
}

}  // namespace

extern "C" int wf_tb_tile_shapes(int32_t *out, int32_t cap) {
    const int n = (int)(sizeof(kTileShapes) / sizeof(kTileShapes[0]));
    if (!out || cap < 0) {
        wf::set_error("wf_tb_tile_shapes: bad arguments");
        return WF_ERR_ARG;
    }
    for (int k = 0; k < n && k < cap; ++k) {
        out[2 * k] = kTileShapes[k][0];
        out[2 * k + 1] = kTileShapes[k][1];  // tile_j is 64*cpl - 2*depth
    }
    return n;
}

extern "C" int wf_fdtd2d_tb_advance(const wf_tb_args *args, hipStream_t stream) {
    if (!args || !args->runs || !args->tiles || args->n_runs < 1 || args->n_tiles < 1) {
        wf::set_error("wf_fdtd2d_tb_advance: bad arguments");
        return WF_ERR_ARG;
    }
    const wf_tb_args &a = *args;
    if (a.depth < 1 || a.depth > WF_TB_MAX_DEPTH) {
        wf::set_error("wf_fdtd2d_tb_advance: depth %d outside [1, %d]", a.depth, WF_TB_MAX_DEPTH);
        return WF_ERR_ARG;
    }
    const int cpl = a.tile_j == wf_tb::kW - 2 * a.depth ? 1 : (a.tile_j == 2 * wf_tb::kW - 2 * a.depth ? 2 : 0);
    if (cpl == 0) {
        wf::set_error("wf_fdtd2d_tb_advance: tile_j must be 64*cpl - 2*depth (%d or %d), got %d",
                      wf_tb::kW - 2 * a.depth, 2 * wf_tb::kW - 2 * a.depth, a.tile_j);
        return WF_ERR_ARG;
    }
    if (a.parity < 0 || a.parity > 3 || (a.parity > 1 && a.hist_stride != 0)) {
        wf::set_error("wf_fdtd2d_tb_advance: parity must be 0..3 (2 and 3 not with a state history)");
        return WF_ERR_ARG;
    }
    if (a.tile_i != 16 && a.tile_i != 32 && a.tile_i != 48 && ((a.tile_i != 64 && a.tile_i != 8) || cpl != 1)) {
        wf::set_error("wf_fdtd2d_tb_advance: unsupported tile_i %d with %d column(s) per lane", a.tile_i, cpl);
        return WF_ERR_UNSUPPORTED;
    }
    if (a.hist_stride < 0 || (a.hist_stride > 0 && (!a.has_materials || a.fast_div != WF_DIV_IEEE || cpl != 1 ||
                                                      (a.tile_i != 32 && a.tile_i != 48) || a.maxabs_out || a.stop))) {
        wf::set_error("wf_fdtd2d_tb_advance: a state history (hist_stride %lld) needs materials, IEEE division, "
                      "tile_i 32 or 48 with one column per lane and no maxabs_out or stop flags", (long long)a.hist_stride);
        return WF_ERR_UNSUPPORTED;
    }
    if (a.tile_lists && (!a.grp_list || !a.det_list)) {
        wf::set_error("wf_fdtd2d_tb_advance: tile_lists needs grp_list and det_list");
        return WF_ERR_ARG;
    }
    if (a.fast_div < WF_DIV_IEEE || a.fast_div > WF_DIV_RECIPROCAL) {
        wf::set_error("wf_fdtd2d_tb_advance: fast_div must be WF_DIV_IEEE, WF_DIV_FAST or WF_DIV_RECIPROCAL, got %d",
                      a.fast_div);
        return WF_ERR_ARG;
    }
    if (a.fast_div == WF_DIV_FAST && (a.div_lo_x != 0u || a.div_lo_y != 0u || a.div_span_x > 0x7f800000u ||
                                      a.div_span_y > 0x7f800000u)) {
        wf::set_error("wf_fdtd2d_tb_advance: WF_DIV_FAST needs div_lo = 0 and div_span <= bits(inf)");
        return WF_ERR_ARG;
    }
    hipError_t e;
    const int dv = a.fast_div;
    if (a.hist_stride > 0)
        e = wf_tb_launch_mat_ieee_hist(a, stream);
    else if (cpl == 2)
        e = a.has_materials ? (dv == WF_DIV_FAST ? wf_tb_launch_mat_fast_w2(a, stream)
                                                 : dv == WF_DIV_RECIPROCAL ? wf_tb_launch_mat_rcp_w2(a, stream)
                                                                           : wf_tb_launch_mat_ieee_w2(a, stream))
                            : (dv == WF_DIV_FAST ? wf_tb_launch_free_fast_w2(a, stream)
                                                 : dv == WF_DIV_RECIPROCAL ? wf_tb_launch_free_rcp_w2(a, stream)
                                                                           : wf_tb_launch_free_ieee_w2(a, stream));
    else if (a.has_materials)
        e = dv == WF_DIV_FAST ? wf_tb_launch_mat_fast(a, stream)
                              : dv == WF_DIV_RECIPROCAL ? wf_tb_launch_mat_rcp(a, stream) : wf_tb_launch_mat_ieee(a, stream);
    else
        e = dv == WF_DIV_FAST ? wf_tb_launch_free_fast(a, stream)
                              : dv == WF_DIV_RECIPROCAL ? wf_tb_launch_free_rcp(a, stream) : wf_tb_launch_free_ieee(a, stream);
    return wf::check_hip(e, "wf_fdtd2d_tb_advance");
}

namespace {

__global__ void k_judge(const unsigned int *__restrict__ maxabs, int n_runs, double threshold, int all,
                        int *__restrict__ stop) {
    __shared__ int s_any;
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    for (int r = threadIdx.x; r < n_runs; r += blockDim.x) {
        bool bad = false;
        for (int c = 0; c < 3; ++c) bad = bad || ((double)__uint_as_float(maxabs[3 * r + c]) > threshold);
        if (bad) {
            stop[r] = 1;
            s_any = 1;
        }
    }
    __syncthreads();
    if (all && s_any)
        for (int r = threadIdx.x; r < n_runs; r += blockDim.x) stop[r] = 1;
}

}  // namespace

extern "C" int wf_fdtd2d_judge(const uint32_t *maxabs, int32_t n_runs, double threshold, int32_t all, int32_t *stop,
                               hipStream_t stream) {
    if (!maxabs || !stop || n_runs < 1) {
        wf::set_error("wf_fdtd2d_judge: bad arguments");
        return WF_ERR_ARG;
    }
    hipLaunchKernelGGL(k_judge, dim3(1), dim3(256), 0, stream, reinterpret_cast<const unsigned int *>(maxabs),
                       n_runs, threshold, all, reinterpret_cast<int *>(stop));
    return wf::check_launch("wf_fdtd2d_judge");
}

namespace {
__global__ void k_add_step(wf_run2d *runs, int n_runs, int delta) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n_runs; r += gridDim.x * blockDim.x) runs[r].step += delta;
}
}  // namespace

extern "C" int wf_runs_add_step(wf_run2d *runs, int32_t n_runs, int32_t delta, hipStream_t stream) {
    if (!runs || n_runs < 1) {
        wf::set_error("wf_runs_add_step: bad arguments");
        return WF_ERR_ARG;
    }
    hipLaunchKernelGGL(k_add_step, dim3((n_runs + 255) / 256), dim3(256), 0, stream, runs, n_runs, delta);
    return wf::check_launch("wf_runs_add_step");
}

extern "C" int wf_div_verify(float d, float r, uint32_t lo, uint32_t span, uint32_t *bad, hipStream_t stream) {
    const unsigned int be = (__builtin_bit_cast(uint32_t, d) >> 23) & 0xffu;
    if (!bad || !(d > 0.0f) || be < 32u || be > 158u || span == 0u || lo + span < lo || lo + span > 0x7f800000u) {
        wf::set_error("wf_div_verify: bad arguments (d must be normal with exponent in [-95, 31], "
                      "0 < span, lo + span <= bits(inf))");
        return WF_ERR_ARG;
    }
    hipLaunchKernelGGL(k_div_verify, dim3(8192), dim3(256), 0, stream, d, r, lo, span, reinterpret_cast<unsigned int *>(bad));
    return wf::check_launch("wf_div_verify");
}
