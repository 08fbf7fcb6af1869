// Temporally blocked, multi-run 2D TM FDTD advance (the throughput path).
//
// Work decomposition. One workgroup of 8 wavefronts owns one core tile of one run: TI rows x
// TJ = W - 2D columns, plus a halo of D cells per side. Every row of the
// staged region is W = 64*CPL cells, one wavefront with CPL adjacent columns
// per lane (lane l holds columns CPL*l .. CPL*l + CPL-1). Wave w holds the
// NR = ceil((TI + 2D)/8) consecutive rows w*NR .. w*NR + NR-1 of the region
// in registers (Ex, Ey, Hz and, on the material path, Ca/Cb). The workgroup
// advances the region D leapfrog steps without touching HBM, shrinking the
// valid region by one cell per side per step, then writes the core to
// another state buffer (no workgroup reads what another writes in the same
// launch).
// HBM traffic is ~(20*R + 12)/D bytes per cell-update (R = region/core area)
// instead of the 32 B of one step per pass. Two columns per lane halve the
// share of halo columns (a 128-wide region keeps 112 columns at D = 8, a
// 64-wide one 48) at twice the registers per staged row.
//
// Neighbours: the j-neighbours of a cell are the lane's other column or the
// adjacent lane's, fetched with a DPP whole-wave shift; the i-neighbours are
// the lane's own registers, except across the boundary between two waves'
// row blocks, where one row per wave goes through LDS each half-step (the Ey
// of its first row for the wave above, the Hz of its last row for the wave
// below). Every tile steps every row of the region without row tests (a
// row outside the valid trapezoid or the grid only feeds rows outside it, or
// the face cells the Mur pass and the kept Ex/Ey columns overwrite); edge
// tiles also keep a copy of every Hz row in LDS for the Mur pass.
//
// Each step reproduces the reference semantics exactly, in the same fp32
// operation order (the library is built with -ffp-contract=off):
//   1. soft sources inside the region, in collection order, added to the
//      owning lane's registers (src/core/sources.py:1060-1071); LDS keeps the
//      pre-injection Hz, which is the Mur snapshot (src/core/boundaries.py:237);
//   2. Faraday on the grid rows 1..nx-2 of the computed range, every lane
//      (src/core/fdtd2d.py:232-235; values on face columns and outside the grid
//      are discarded); on tiles that touch the grid boundary, a short pass
//      after it (one thread per face cell, from the Faraday values and the
//      snapshot of the face rows and columns and those next to them, kept in
//      LDS) writes the Mur faces (src/core/boundaries.py:259-274), and the rows
//      are reloaded from LDS; interior tiles skip both;
//   3. Ampere (src/core/fdtd2d.py:244-257) on every lane and row of the
//      computed range; the cells the reference never writes (Ex of column
//      j = 0, Ey of row i = 0) are saved before it and put back after it on
//      the tiles that hold them, so interior tiles pay no per-cell select;
//      Hz detectors owned by the core are read from LDS (a tile that holds
//      detectors keeps every Hz row there)
//      (examples/brain_mimo_imaging.py:196-198).
//
// Division by the cell sizes, DIV (wf_tb_args.fast_div):
//   kDivIeee: IEEE x/fp32(d), as torch on the CPU divides by a Python float;
//   kDivFast: the reciprocal product corrected by one fused residual step
//     taken at a scale s = 2^k that puts d*s in [2^31, 2^32): q0 = x*r,
//     t = q0*(d*s) - x*s, q = q0 - t*(r/s). The scaling keeps the residual
//     exact down to the smallest subnormal x, and the operand signs make q = x
//     for x = +0 and x = -0. The library proves q equal to IEEE x/d for every
//     |x| below a bound h (wf_div_verify, exhaustive per divisor); NaN and
//     |x| >= h fall back to IEEE division;
//   kDivRcp: x*fp32(1/d), the reciprocal of the double cell size rounded once
//     (each run's rdx, rdy), as torch on a GPU divides by a Python float: one
//     product per quotient instead of five instructions.
//
// Guards. A source step past the end of its waveform table injects nothing
// and sets WF_STATUS_WAVE_END in the run's status word; a tile region with
// more source groups, or a tile core with more detectors, than the kernel
// stages sets WF_STATUS_TILE_GROUPS / WF_STATUS_TILE_DETS (the extra ones are
// not applied). wf_status_check reports both as errors.
//
// Tiles are mapped XCD-aware: the dispatcher deals workgroups round-robin to
// the 8 XCDs, so block b runs on XCD b%8; we hand each XCD a contiguous band
// of tiles so neighbouring halos are read through the same L2.
#pragma once
#include "wf_common.hpp"
#include "wf_div.hpp"


namespace wf_tb {

constexpr int kW = 64;          // lanes per wavefront
constexpr int kMaxGroups = 64;  // source groups per tile region (host checks)
constexpr int kMaxDets = 64;    // detectors per tile core (host checks)

__device__ __forceinline__ float lane_next(float v) {  // value of lane+1 (0 in lane 63)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float v) {  // value of lane-1 (0 in lane 0)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// One grid row of one field as a raw buffer: a wave-uniform descriptor and a
// 32-bit byte offset per lane. An offset at or past the row's end (kOut, or
// any offset when the row is outside the grid and has 0 bytes) loads 0 and
// stores nothing, so the loads and stores need no per-lane address
// arithmetic or masks.
constexpr uint32_t kOut = 0xffffffffu;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_of(const float *base, int64_t o, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base + o), 0, bytes, 0x00020000);
}
__device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t p, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p, off, 0, 0));
}
__device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t p, uint32_t off, float v) {
    // default cache policy: the next launch reads these cells (the
    // non-temporal hint measured 0-9 % slower, most at shallow depths)
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), p, off, 0, 0);
}
// A pointer the whole wavefront shares, kept in scalar registers.
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

using wf::Divider;
using wf::divide;
using wf::divide_row;
using wf::kDivFast;
using wf::kDivIeee;
using wf::kDivRcp;
using wf::fast_quotient;
using wf::make_divider;

// CPL consecutive floats of one LDS row at a lane's first column (8-byte
// accesses for two columns per lane).
template <int CPL>
__device__ __forceinline__ void lds_get(const float *p, float (&v)[CPL]) {
    if constexpr (CPL == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        v[0] = t.x;
        v[1] = t.y;
    } else {
        v[0] = p[0];
    }
}
template <int CPL>
__device__ __forceinline__ void lds_put(float *p, const float (&v)[CPL]) {
    if constexpr (CPL == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
    } else {
        p[0] = v[0];
    }
}

// Wavefronts per workgroup, rows per lane and the occupancy the register
// budget is sized for. Every shape runs 8 wavefronts: the more waves share a
// region, the fewer rows each walks per step, which is the latency of a
// launch on grids too small to fill the GPU (a 16-row tile holds 4 rows per
// lane). Each lane keeps 5*CPL values per row (Ex, Ey, Hz, Ca, Cb) plus the
// temporaries of its division mode (36 registers for reciprocal products, 40
// for the fast quotient or IEEE division); without the cap the scheduler
// hoists every row's loads and runs at one or two waves per SIMD. An
// 8-wavefront workgroup puts two waves on each SIMD, so its occupancy is an
// even number of waves per SIMD: 8 (64 registers), 6 (80) or 4 (128).
template <int TI, int CPL>
constexpr int kWavesOf = 8;
template <int TI, int CPL>
constexpr int kThreadsOf = kWavesOf<TI, CPL> * kW;

template <int TI, int D, int CPL, int DIV = kDivRcp>
struct TileGeom {
    static constexpr int NW = kWavesOf<TI, CPL>;
    static constexpr int NT = NW * kW;
    static constexpr int W = kW * CPL;
    static constexpr int RI = TI + 2 * D;
    static constexpr int NR = (RI + NW - 1) / NW;  // wave w holds rows w*NR .. w*NR + NR-1
    static constexpr int kRegs = 5 * NR * CPL + (DIV == kDivRcp ? 36 : 40);
    static constexpr int kWaves = kRegs <= 64 ? 8 : (kRegs <= 80 ? 6 : (kRegs <= 128 ? 4 : 2));
    static_assert(NR <= 32, "row flags are one 32-bit word per wavefront");
    static_assert(NT >= 2 * W + RI, "edge passes need a thread per face cell");
};

// HIST: write the core after every step s to the written state + (s-1)*hist_stride
// (a state history for the adjoint's recompute) instead of only after the last.
template <int TI, int D, bool MAT, int DIV, bool HIST = false, int CPL = 1>
__global__ void __launch_bounds__((kThreadsOf<TI, CPL>)) __attribute__((amdgpu_waves_per_eu((TileGeom<TI, D, CPL, DIV>::kWaves))))
k_tb(wf_tb_args a) {
    using G = TileGeom<TI, D, CPL, DIV>;
    constexpr int W = G::W;
    constexpr int TJ = W - 2 * D;
    constexpr int RI = G::RI;
    constexpr int NW = G::NW;
    constexpr int kNT = G::NT;
    constexpr int NR = G::NR;
    __shared__ __align__(16) float s_hz[RI * W];  // edge tiles and tiles with detectors: every row's Hz
    // row exchange between the waves' row blocks, one array so every access is
    // one base register plus a constant: [0, XH) holds at entry w the Ey of
    // wave w's first row (read by wave w-1), [XH, 2 XH) at entry w+1 the Hz of
    // wave w's last row (read by wave w+1)
    constexpr int XH = (NW + 1) * W;
    __shared__ __align__(16) float s_x[2 * XH];
    __shared__ int s_grp[kMaxGroups];
    __shared__ int s_grp_cell[kMaxGroups];
    __shared__ int s_det[kMaxDets];
    __shared__ int s_det_cell[kMaxDets];
    __shared__ int s_ngrp, s_ndet;
    __shared__ unsigned int s_rowmask[NW];  // per wavefront: bit k set if its row w*NR+k holds a source
    __shared__ unsigned int s_max[3];
    __shared__ float s_prow[2][W];   // edge tiles: snapshot of rows i = 1 and i = nx-2
    __shared__ float s_frow[2][W];   // edge tiles: snapshot of the face rows i = 0 and i = nx-1
    __shared__ float s_pcol[2][RI];  // edge tiles: snapshot of columns j = 1 and j = ny-2
    __shared__ float s_fcol[2][RI];  // edge tiles: snapshot of the face columns j = 0 and j = ny-1
    __shared__ float s_keep_ex[RI];  // left tiles: Ex of column j = 0 before Ampere
    __shared__ __align__(16) float s_keep_ey[W];  // top tiles: Ey of row i = 0 before Ampere

    const int tid = threadIdx.x;
    const int lane = tid & (kW - 1);
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
    const int per = (a.n_tiles + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= a.n_tiles) return;
    const int run = a.tiles[2 * t];
    const int tile = a.tiles[2 * t + 1];
    const wf_run2d &R = a.runs[run];
    const int nx = R.nx, ny = R.ny;
    const int ntj = (ny + TJ - 1) / TJ;
    const int I0 = (tile / ntj) * TI;
    const int J0 = (tile % ntj) * TJ;
    const int gi0 = I0 - D, gj0 = J0 - D;  // global coordinates of region (0, 0)
    const int c0 = lane * CPL;             // region column of the lane's first cell
    const int r0 = w * NR;                 // region row of the wave's first row

    // parity 0: A -> B, 1: B -> A, 2: A -> C, 3: C -> B (A = ex/ey/hz, B = ex2.., C = ex3..)
    const int par = a.parity;
    const float *src_ex = par == 1 ? R.ex2 : (par == 3 ? R.ex3 : R.ex);
    const float *src_ey = par == 1 ? R.ey2 : (par == 3 ? R.ey3 : R.ey);
    const float *src_hz = par == 1 ? R.hz2 : (par == 3 ? R.hz3 : R.hz);
    float *dst_ex = par == 1 ? R.ex : (par == 2 ? R.ex3 : R.ex2);
    float *dst_ey = par == 1 ? R.ey : (par == 2 ? R.ey3 : R.ey2);
    float *dst_hz = par == 1 ? R.hz : (par == 2 ? R.hz3 : R.hz2);

    // the step index, read once: the kernel's own stores (traces) could alias the
    // descriptor as far as the compiler knows, so a read inside the loop would
    // be a global load and wait on every step
    const int n0 = __builtin_amdgcn_readfirstlane(R.step + a.step_offset);
    const int n_end = __builtin_amdgcn_readfirstlane(R.step_end > 0 ? R.step_end : 0x7fffffff);
    if ((a.stop != nullptr && a.stop[run] != 0) || n0 >= n_end) {
        // The run failed a stability check earlier in this stream, or has
        // reached its last step: carry its state through unchanged (so it ends
        // in the live buffers as it was), folding its maxima if asked.
        unsigned int m0 = 0u, m1 = 0u, m2 = 0u;
        for (int k = w; k < TI; k += NW) {
            const int gi = I0 + k;
            for (int c = lane; c < TJ; c += kW) {
                const int cj = J0 + c;
                if (gi < nx && cj < ny) {
                    const size_t g = (size_t)gi * ny + cj;
                    const float vx = src_ex[g], vy = src_ey[g], vh = src_hz[g];
                    dst_ex[g] = vx;
                    dst_ey[g] = vy;
                    dst_hz[g] = vh;
                    m0 = max(m0, wf::abs_bits(vx));
                    m1 = max(m1, wf::abs_bits(vy));
                    m2 = max(m2, wf::abs_bits(vh));
                }
            }
        }
        if (a.maxabs_out) {
            for (int off = 32; off > 0; off >>= 1) {
                m0 = max(m0, (unsigned int)__shfl_xor((int)m0, off));
                m1 = max(m1, (unsigned int)__shfl_xor((int)m1, off));
                m2 = max(m2, (unsigned int)__shfl_xor((int)m2, off));
            }
            if (lane == 0) {
                atomicMax(&a.maxabs_out[3 * run + 0], m0);
                atomicMax(&a.maxabs_out[3 * run + 1], m1);
                atomicMax(&a.maxabs_out[3 * run + 2], m2);
            }
        }
        return;
    }

    if (tid == 0) {
        s_ngrp = 0;
        s_ndet = 0;
        s_max[0] = s_max[1] = s_max[2] = 0u;
    }
    if (tid < NW) s_rowmask[tid] = 0u;

    bool col_in[CPL];
#pragma unroll
    for (int h = 0; h < CPL; ++h) {
        const int gj = gj0 + c0 + h;
        col_in[h] = gj >= 0 && gj < ny;
    }
    // the wave's NR consecutive rows (rows past the region and outside the grid hold 0)
    uint32_t coff[CPL];  // byte offset of the lane's cells in a row, kOut outside the grid's columns
#pragma unroll
    for (int h = 0; h < CPL; ++h) coff[h] = col_in[h] ? 4u * (uint32_t)(gj0 + c0 + h) : kOut;
    // the tile's row of the per-tile lists, in flight with the field loads
    // (read after the barrier below, it would add a load's latency to every
    // workgroup before its first step)
    int4 tl = make_int4(0, R.n_grp, 0, R.n_det);
    if (a.tile_lists) tl = reinterpret_cast<const int4 *>(a.tile_lists)[t];
    const float *p_ex = uniform_ptr(src_ex), *p_ey = uniform_ptr(src_ey), *p_hz = uniform_ptr(src_hz);
    const float *p_ca = MAT ? uniform_ptr(R.ca) : nullptr, *p_cb = MAT ? uniform_ptr(R.cb) : nullptr;
    float ex[NR][CPL], ey[NR][CPL], hz[NR][CPL], ca[NR][CPL], cb[NR][CPL];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int gi = gi0 + r0 + k;
        const bool in = r0 + k < RI && gi >= 0 && gi < nx;
        const int64_t ro = in ? (int64_t)gi * ny : 0;
        const int nb = in ? 4 * ny : 0;
        const auto rx = row_of(p_ex, ro, nb), ry = row_of(p_ey, ro, nb), rh = row_of(p_hz, ro, nb);
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            ex[k][h] = ld(rx, coff[h]);
            ey[k][h] = ld(ry, coff[h]);
            hz[k][h] = ld(rh, coff[h]);
            ca[k][h] = cb[k][h] = 0.f;
        }
        if (MAT) {
            const auto ra = row_of(p_ca, ro, nb), rb = row_of(p_cb, ro, nb);
#pragma unroll
            for (int h = 0; h < CPL; ++h) {
                ca[k][h] = ld(ra, coff[h]);
                cb[k][h] = ld(rb, coff[h]);
            }
        }
    }
    __syncthreads();
    // this tile's source groups and detectors: from the host's per-tile lists
    // (their row read with the fields, above), or by testing every group and
    // detector of the run
    const int g_lo = tl.x, g_hi = tl.y, d_lo = tl.z, d_hi = tl.w;
    for (int q = g_lo + tid; q < g_hi; q += kNT) {
        const int g = a.tile_lists ? a.grp_list[q] : q;
        const wf_inj2d e0 = R.inj[R.grp_start[g]];
        const int li = e0.cell / ny - gi0, lj = e0.cell % ny - gj0;
        if (li >= 0 && li < RI && lj >= 0 && lj < W) {
            const int k = atomicAdd(&s_ngrp, 1);
            if (k < kMaxGroups) {
                s_grp[k] = g;
                s_grp_cell[k] = li * W + lj;
                atomicOr(&s_rowmask[li / NR], 1u << (li % NR));
            } else if (R.status) {
                atomicOr(R.status, (unsigned int)WF_STATUS_TILE_GROUPS);
            }
        }
    }
    for (int q = d_lo + tid; q < d_hi; q += kNT) {
        const int d = a.tile_lists ? a.det_list[q] : q;
        const int cell = R.det_cell[d];
        const int gi = cell / ny, gjd = cell % ny;
        if (gi >= I0 && gi < I0 + TI && gjd >= J0 && gjd < J0 + TJ) {
            const int k = atomicAdd(&s_ndet, 1);
            if (k < kMaxDets) {
                s_det[k] = d;
                s_det_cell[k] = (gi - gi0) * W + (gjd - gj0);
            } else if (R.status) {
                atomicOr(R.status, (unsigned int)WF_STATUS_TILE_DETS);
            }
        }
    }
    __syncthreads();
    const int ngrp = min(s_ngrp, kMaxGroups);
    const int ndet = min(s_ndet, kMaxDets);
    const unsigned int rowmask = s_rowmask[w];
    // Edge tiles (touching a grid face) keep every Hz row in LDS, which the
    // Mur pass reads and writes; interior tiles keep Hz rows in LDS only when
    // they record detectors. Both step every row of the region.
    const bool edge = (gi0 <= 0) || (gi0 + RI >= nx) || (gj0 <= 0) || (gj0 + W >= ny);
    const bool mirror = edge || ndet > 0;
    const float dh = R.dh, de = R.de, cx = R.cmx, cy = R.cmy;
    Divider divx{}, divy{};
    if constexpr (DIV == kDivRcp) {  // quotients are products by the run's fp32(1/d)
        divx.r = R.rdx;
        divy.r = R.rdy;
    } else {
        divx = make_divider(R.dx, a.rdx, a.div_span_x);
        divy = make_divider(R.dy, a.rdy, a.div_span_y);
    }
    (void)de;
    bool lane_core[CPL];
#pragma unroll
    for (int h = 0; h < CPL; ++h) lane_core[h] = c0 + h >= D && c0 + h < D + TJ && col_in[h];
    uint32_t soff[CPL];  // byte offset of the lane's core cells in a row, kOut elsewhere
#pragma unroll
    for (int h = 0; h < CPL; ++h) soff[h] = lane_core[h] ? 4u * (uint32_t)(gj0 + c0 + h) : kOut;
    float *q_ex = uniform_ptr(dst_ex), *q_ey = uniform_ptr(dst_ey), *q_hz = uniform_ptr(dst_hz);

    // Faraday at row k from its registers and the Ey row below (eyd):
    // Hz += Dh*((Ex(j+1) - Ex(j))/dy - (Ey(i+1) - Ey(i))/dx), every lane (on
    // the face columns and outside the grid the value is discarded or feeds
    // only discarded values)
    auto faraday = [&](const int k, const float (&eyd)[CPL]) __attribute__((always_inline)) {
        float x1[CPL], x2[CPL], av[CPL], bv[CPL];
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            const float exn = h + 1 < CPL ? ex[k][h + 1] : lane_next(ex[k][0]);
            x1[h] = exn - ex[k][h];
            x2[h] = eyd[h] - ey[k][h];
        }
        divide_row<DIV, CPL>(x1, divy, x2, divx, av, bv);
#pragma unroll
        for (int h = 0; h < CPL; ++h) hz[k][h] = hz[k][h] + dh * (av[h] - bv[h]);
    };
    // Ampere at row k from its registers and the (new) Hz row above (hzu)
    auto ampere = [&](const int k, const float (&hzu)[CPL]) __attribute__((always_inline)) {
        float x3[CPL], x4[CPL], gyv[CPL], gxv[CPL];
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            const float hzl = h > 0 ? hz[k][h - 1] : lane_prev(hz[k][CPL - 1]);
            x3[h] = hz[k][h] - hzl;
            x4[h] = hz[k][h] - hzu[h];
        }
        divide_row<DIV, CPL>(x3, divy, x4, divx, gyv, gxv);
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            if (MAT) {
                const float t1 = ca[k][h] * ex[k][h];
                const float t2 = cb[k][h] * gyv[h];
                ex[k][h] = t1 + t2;
                const float u1 = ca[k][h] * ey[k][h];
                const float u2 = cb[k][h] * gxv[h];
                ey[k][h] = u1 - u2;
            } else {
                const float t2 = de * gyv[h];
                ex[k][h] = ex[k][h] + t2;
                const float u2 = de * gxv[h];
                ey[k][h] = ey[k][h] - u2;
            }
        }
    };

    // Edge snapshot copies (before the first step, and after each step's Mur
    // pass for the next): rows i = 1, nx-2 and the face rows i = 0, nx-1;
    // columns j = 1, ny-2 and the face columns j = 0, ny-1. LDS holds the
    // state every wave's registers hold, before any source of the step.
    auto edge_snapshot = [&]() __attribute__((always_inline)) {
        // the edge passes' values are formed inside their branches (the
        // empty asm keeps the compiler from hoisting them out of the step
        // loop, where they would hold registers for every tile)
        int e_i0 = gi0, e_j0 = gj0, e_t = tid;
        asm volatile("" : "+s"(e_i0), "+s"(e_j0), "+v"(e_t));
        const int r1 = 1 - e_i0, r2 = nx - 2 - e_i0;  // region rows of i = 1, nx-2
        const int q1 = 1 - e_j0, q2 = ny - 2 - e_j0;  // region columns of j = 1, ny-2
        if (e_t < W) {
            if (r1 >= 0 && r1 < RI) s_prow[0][e_t] = s_hz[r1 * W + e_t];
            if (r2 >= 0 && r2 < RI) s_prow[1][e_t] = s_hz[r2 * W + e_t];
            if (r1 >= 1 && r1 <= RI) s_frow[0][e_t] = s_hz[(r1 - 1) * W + e_t];
            if (r2 >= -1 && r2 + 1 < RI) s_frow[1][e_t] = s_hz[(r2 + 1) * W + e_t];
        } else if (e_t >= 2 * W && e_t < 2 * W + RI) {
            const int li = e_t - 2 * W;
            if (q1 >= 0 && q1 < W) s_pcol[0][li] = s_hz[li * W + q1];
            if (q2 >= 0 && q2 < W) s_pcol[1][li] = s_hz[li * W + q2];
            if (q1 >= 1 && q1 <= W) s_fcol[0][li] = s_hz[li * W + q1 - 1];
            if (q2 >= -1 && q2 + 1 < W) s_fcol[1][li] = s_hz[li * W + q2 + 1];
        }
    };
    float *const px = &s_x[w * W + c0];       // the wave's exchange entries
    float *const hrow = &s_hz[r0 * W + c0];   // the wave's first Hz row in LDS
    if (edge) {  // the first step's Mur snapshot
#pragma unroll
        for (int k = 0; k < NR; ++k)
            if (r0 + k < RI) lds_put<CPL>(&hrow[k * W], hz[k]);
        __syncthreads();
        edge_snapshot();
    }
#pragma unroll 1
    for (int s = 1; s <= D; ++s) {
        const int n = n0 + s - 1;
        if (n >= n_end) break;  // the run's last step was taken: its state stays as it is
        // 1. Soft sources into the owning lane's registers (rows that hold a
        // source are flagged per wavefront, so the scan is skipped elsewhere).
        if (ngrp > 0) {
            // tested here, not hoisted out of the step loop as one held mask per row
            unsigned int rm = __builtin_amdgcn_readfirstlane(rowmask);  // one per wavefront
            asm volatile("" : "+s"(rm));
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (rm & (1u << k)) {
                    const int li = r0 + k;
                    for (int q = 0; q < ngrp; ++q) {
                        const int cl = s_grp_cell[q];
#pragma unroll
                        for (int h = 0; h < CPL; ++h) {
                            if (cl == li * W + c0 + h) {
                                const int g = s_grp[q];
                                const int b0 = R.grp_start[g], b1 = R.grp_start[g + 1];
                                const int comp = R.inj[b0].comp;
                                float v = comp == WF_COMP_EX ? ex[k][h] : (comp == WF_COMP_EY ? ey[k][h] : hz[k][h]);
                                for (int e = b0; e < b1; ++e) {
                                    const wf_inj2d &inj = R.inj[e];
                                    if ((unsigned)n < (unsigned)inj.wave_len) {
                                        const float wv = inj.wave[n] * inj.amp;
                                        v = v + wv;
                                    } else if (R.status) {
                                        atomicOr(R.status, (unsigned int)WF_STATUS_WAVE_END);
                                    }
                                }
                                // value selects, not a store through a selected array
                                // (which would move the rows to scratch memory)
                                ex[k][h] = comp == WF_COMP_EX ? v : ex[k][h];
                                ey[k][h] = comp == WF_COMP_EY ? v : ey[k][h];
                                hz[k][h] = comp == WF_COMP_HZ ? v : hz[k][h];
                            }
                        }
                    }
                }
            }
        }
        // 2. The Ey of each wave's first row, which the wave above reads for
        // its last row's Faraday (after the sources: Faraday reads injected Ey).
        lds_put<CPL>(&px[0], ey[0]);
        __syncthreads();
        // 3. Faraday on every row; its values on the face rows and columns
        // are replaced by the Mur pass below, and rows outside the valid
        // trapezoid feed only rows outside it.
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            float eyd[CPL];
            if (k + 1 < NR) {
#pragma unroll
                for (int h = 0; h < CPL; ++h) eyd[h] = ey[k + 1][h];
            } else {
                lds_get<CPL>(&px[W], eyd);
            }
            faraday(k, eyd);
        }
        if (mirror) {  // edge tiles (the Mur pass) and detectors read Hz rows from LDS
#pragma unroll
            for (int k = 0; k < NR; ++k)
                if (r0 + k < RI) lds_put<CPL>(&hrow[k * W], hz[k]);
        }
        lds_put<CPL>(&px[XH + W], hz[NR - 1]);
        __syncthreads();
        if (edge) {
            // Mur (src/core/boundaries.py:259-274) on the faces next to this
            // step's valid Faraday rows [s-1, RI-s): f[face] = P[in] + C (f[in] - P[face]),
            // x faces first; a corner takes the y-face formula over the x-face
            // value beside it. One thread per face cell; every value it reads
            // is Faraday's or the snapshot's, so the pass needs no barrier.
            int e_i0 = gi0, e_j0 = gj0, e_t = tid;
            asm volatile("" : "+s"(e_i0), "+s"(e_j0), "+v"(e_t));
            const int r1 = 1 - e_i0, r2 = nx - 2 - e_i0;
            const int q1 = 1 - e_j0, q2 = ny - 2 - e_j0;
            if (e_t < 2 * W) {  // x faces: rows i = 0 (from i = 1) and i = nx-1 (from i = nx-2)
                const int side = e_t >= W, l = e_t - side * W;
                const int rin = side ? r2 : r1, rf = side ? rin + 1 : rin - 1;
                const int gjl = e_j0 + l;
                if (rin >= s - 1 && rin < RI - s && rf >= 0 && rf < RI && gjl >= 1 && gjl <= ny - 2) {
                    const float p_in = s_prow[side][l], p_f = s_frow[side][l];
                    const float L = p_in + cx * (s_hz[rin * W + l] - p_f);
                    if (gjl == 1 && l >= 1) {
                        const float pc = s_frow[side][l - 1];
                        s_hz[rf * W + l - 1] = p_f + cy * (L - pc);
                    }
                    if (gjl == ny - 2 && l + 1 < W) {
                        const float pc = s_frow[side][l + 1];
                        s_hz[rf * W + l + 1] = p_f + cy * (L - pc);
                    }
                    s_hz[rf * W + l] = L;
                }
            } else if (e_t - 2 * W < RI) {  // y faces: j = 0 (from j = 1) and j = ny-1 (from j = ny-2)
                const int li = e_t - 2 * W, gil = e_i0 + li;
                if (li >= s - 1 && li < RI - s && gil >= 1 && gil <= nx - 2) {
                    if (q1 >= 1 && q1 < W) {
                        const float pf = s_fcol[0][li];
                        s_hz[li * W + q1 - 1] = s_pcol[0][li] + cy * (s_hz[li * W + q1] - pf);
                    }
                    if (q2 >= 0 && q2 + 1 < W) {
                        const float pf = s_fcol[1][li];
                        s_hz[li * W + q2 + 1] = s_pcol[1][li] + cy * (s_hz[li * W + q2] - pf);
                    }
                }
            }
            __syncthreads();
            // face cells take their Mur values: LDS equals the registers on
            // every other cell, so whole rows are reloaded
#pragma unroll
            for (int k = 0; k < NR; ++k)
                if (r0 + k < RI) lds_get<CPL>(&hrow[k * W], hz[k]);
            // the next step's snapshot: LDS holds this step's final Hz, which
            // nothing rewrites before the next Faraday (Ampere and the
            // detectors only read it; the next step's first barrier orders
            // these copies before it)
            if (s < D) edge_snapshot();
            // The cells Ampere must leave as they are: Ex of column j = 0 and
            // Ey of row i = 0 (src/core/fdtd2d.py:244-257 writes Ex[:, 1:] and
            // Ey[1:, :]); kept per lane (LDS written and read by the same
            // lane) and put back after Ampere.
            if (e_j0 <= 0) {
                const int cz = -e_j0;  // region column of j = 0
                if (lane == cz / CPL) {
#pragma unroll
                    for (int k = 0; k < NR; ++k) {
                        const int li = r0 + k;
                        float v = ex[k][0];
#pragma unroll
                        for (int h = 1; h < CPL; ++h) v = cz % CPL == h ? ex[k][h] : v;
                        if (li < RI) s_keep_ex[li] = v;
                    }
                }
            }
            if (e_i0 <= 0 && w == (-e_i0) / NR) {
                const int kz = (-e_i0) % NR;  // wave w holds region row -gi0 (i = 0) as its row kz
                // a select chain, not ey[kz] (a dynamic index would move the rows to scratch)
                float v[CPL];
#pragma unroll
                for (int h = 0; h < CPL; ++h) v[h] = ey[0][h];
#pragma unroll
                for (int k = 1; k < NR; ++k)
#pragma unroll
                    for (int h = 0; h < CPL; ++h) v[h] = k == kz ? ey[k][h] : v[h];
                lds_put<CPL>(&s_keep_ey[c0], v);
            }
        }
        // 4. Detectors (Hz after the step), then Ampere on every row.
        if (tid < ndet) {
            const int tt = n - R.trace_n0;
            if (tt >= 0 && tt < R.trace_len) R.traces[s_det[tid] * R.trace_stride + tt] = s_hz[s_det_cell[tid]];
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            float hzu[CPL];
            if (k > 0) {
#pragma unroll
                for (int h = 0; h < CPL; ++h) hzu[h] = hz[k - 1][h];
            } else if (edge) {  // the row above, after the Mur pass
                lds_get<CPL>(&s_hz[max(r0 - 1, 0) * W + c0], hzu);
            } else {
                lds_get<CPL>(&px[XH], hzu);
            }
            ampere(k, hzu);
        }
        if (edge) {
            int e_i0 = gi0, e_j0 = gj0;
            asm volatile("" : "+s"(e_i0), "+s"(e_j0));
            if (e_j0 <= 0) {
                const int cz = -e_j0;
                if (lane == cz / CPL) {
#pragma unroll
                    for (int k = 0; k < NR; ++k) {
                        const int li = r0 + k;
                        if (li < RI) {
                            const float v = s_keep_ex[li];
#pragma unroll
                            for (int h = 0; h < CPL; ++h)
                                if (cz % CPL == h) ex[k][h] = v;
                        }
                    }
                }
            }
            if (e_i0 <= 0 && w == (-e_i0) / NR) {
                const int kz = (-e_i0) % NR;
                float v[CPL];
                lds_get<CPL>(&s_keep_ey[c0], v);
#pragma unroll
                for (int k = 0; k < NR; ++k)
#pragma unroll
                    for (int h = 0; h < CPL; ++h) ey[k][h] = k == kz ? v[h] : ey[k][h];
            }
        }
        if (HIST) {
            // the state after step s into its history slot (the written state + (s-1)*hist_stride)
            const int64_t hoff = (int64_t)(s - 1) * a.hist_stride;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const int gi = gi0 + r0 + k;
                // wave-uniform: halo rows issue no history stores at all
                if (!(r0 + k >= D && r0 + k < D + TI && gi < nx)) continue;
                const int64_t ro = hoff + (int64_t)gi * ny;
                const int nb = 4 * ny;
                const auto rx = row_of(q_ex, ro, nb), ry = row_of(q_ey, ro, nb), rh = row_of(q_hz, ro, nb);
#pragma unroll
                for (int h = 0; h < CPL; ++h) {
                    st(rx, soff[h], ex[k][h]);
                    st(ry, soff[h], ey[k][h]);
                    st(rh, soff[h], hz[k][h]);
                }
            }
        }
    }
    if (HIST) return;  // the last step's state is in its history slot

    // Write the core back (and fold the stability maxima if requested).
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int li = r0 + k;
        const int gi = gi0 + li;
        const bool core = li >= D && li < D + TI && gi < nx;
        const int64_t ro = core ? (int64_t)gi * ny : 0;
        const int nb = core ? 4 * ny : 0;
        const auto rx = row_of(q_ex, ro, nb), ry = row_of(q_ey, ro, nb), rh = row_of(q_hz, ro, nb);
#pragma unroll
        for (int h = 0; h < CPL; ++h) {
            st(rx, soff[h], ex[k][h]);
            st(ry, soff[h], ey[k][h]);
            st(rh, soff[h], hz[k][h]);
        }
    }
    if (a.maxabs_out) {
        unsigned int m0 = 0u, m1 = 0u, m2 = 0u;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int li = r0 + k;
            const int gi = gi0 + li;
#pragma unroll
            for (int h = 0; h < CPL; ++h) {
                if (li >= D && li < D + TI && gi < nx && lane_core[h]) {
                    m0 = max(m0, wf::abs_bits(ex[k][h]));
                    m1 = max(m1, wf::abs_bits(ey[k][h]));
                    m2 = max(m2, wf::abs_bits(hz[k][h]));
                }
            }
        }
        atomicMax(&s_max[0], m0);
        atomicMax(&s_max[1], m1);
        atomicMax(&s_max[2], m2);
        __syncthreads();
        if (tid < 3) atomicMax(&a.maxabs_out[3 * run + tid], s_max[tid]);
    }
}

template <int TI, bool MAT, int DIV, bool HIST = false, int CPL = 1>
hipError_t launch_depth(const wf_tb_args &a, hipStream_t st) {
    const unsigned blocks = 8u * (unsigned)((a.n_tiles + 7) / 8);
    switch (a.depth) {
#define WF_TB_CASE(DD)                                                                                        \
    case DD:                                                                                                  \
        hipLaunchKernelGGL((k_tb<TI, DD, MAT, DIV, HIST, CPL>), dim3(blocks), dim3(TileGeom<TI, DD, CPL>::NT), 0, \
                           st, a);                                                                            \
        break;
        WF_TB_CASE(1)
        WF_TB_CASE(2)
        WF_TB_CASE(3)
        WF_TB_CASE(4)
        WF_TB_CASE(5)
        WF_TB_CASE(6)
        WF_TB_CASE(7)
        WF_TB_CASE(8)
#undef WF_TB_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace wf_tb

// One object file per (materials, division, columns per lane) instantiates the
// kernels for every tile height and depth; the entry point in wf_tb.hip
// dispatches here. Tile widths: tile_j = 64*CPL - 2*depth.
#define WF_TB_DECLARE(NAME) hipError_t NAME(const wf_tb_args &a, hipStream_t st)
WF_TB_DECLARE(wf_tb_launch_mat_fast);
WF_TB_DECLARE(wf_tb_launch_mat_ieee);
WF_TB_DECLARE(wf_tb_launch_free_fast);
WF_TB_DECLARE(wf_tb_launch_free_ieee);
WF_TB_DECLARE(wf_tb_launch_mat_fast_w2);
WF_TB_DECLARE(wf_tb_launch_mat_ieee_w2);
WF_TB_DECLARE(wf_tb_launch_free_fast_w2);
WF_TB_DECLARE(wf_tb_launch_free_ieee_w2);
WF_TB_DECLARE(wf_tb_launch_mat_rcp);
WF_TB_DECLARE(wf_tb_launch_free_rcp);
WF_TB_DECLARE(wf_tb_launch_mat_rcp_w2);
WF_TB_DECLARE(wf_tb_launch_free_rcp_w2);
WF_TB_DECLARE(wf_tb_launch_mat_ieee_hist);  // hist_stride != 0: 32- or 48-row tiles, one column per lane

namespace wf_tb {
template <bool MAT, int DIV, int CPL>
hipError_t launch_tile(const wf_tb_args &a, hipStream_t st) {
    switch (a.tile_i) {
        case 8:
            if constexpr (CPL == 1) return launch_depth<8, MAT, DIV, false, CPL>(a, st);
            return hipErrorInvalidValue;
        case 16: return launch_depth<16, MAT, DIV, false, CPL>(a, st);
        case 32: return launch_depth<32, MAT, DIV, false, CPL>(a, st);
        case 48: return launch_depth<48, MAT, DIV, false, CPL>(a, st);
        case 64:
            if constexpr (CPL == 1) return launch_depth<64, MAT, DIV, false, CPL>(a, st);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}
}  // namespace wf_tb

#define WF_TB_DEFINE(NAME, MAT, DIV, CPL) \
    WF_TB_DECLARE(NAME) { return wf_tb::launch_tile<MAT, DIV, CPL>(a, st); }
