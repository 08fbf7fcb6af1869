// Temporally blocked, multi-run 2D TM FDTD advance (the throughput path).
//
// One workgroup owns one core tile (TI x TJ cells) of one run. It stages the
// tile plus a halo of D cells per side from HBM into LDS (Ex, Ey, Hz and, on
// the material path, Ca/Cb), advances the whole staged region D leapfrog
// steps in LDS, shrinking the region it computes by one cell per side per
// step, and writes only the core back, into the second state buffer
// (ping-pong: no workgroup reads what another one writes in the same
// launch). HBM traffic per cell-update therefore drops from the 32 B of the
// one-step-per-pass algorithm to about (20*halo_factor + 12)/D bytes.
//
// Each step inside the block reproduces the reference semantics exactly, in
// the same fp32 operation order (the library is built with
// -ffp-contract=off):
//   1. Mur snapshot of the edge rows/columns present in the region
//      (src/core/boundaries.py:230-237) into small LDS strips;
//   2. soft sources falling inside the region, in collection order
//      (src/core/sources.py:1060-1071);
//   3. Faraday on interior cells (src/core/fdtd2d.py:232-235); Mur edge cells
//      (src/core/boundaries.py:259-274) are written by the interior cell whose
//      Faraday value they read, as in the per-step path;
//   4. Ampere (src/core/fdtd2d.py:244-257), Hz detectors owned by the core
//      (examples/brain_mimo_imaging.py:196-198).
// A barrier separates the phases. Cells outside the grid are staged as zero
// and never updated, and no valid cell ever reads them.
//
// Tiles are mapped XCD-aware: the dispatcher deals workgroups round-robin to
// the 8 XCDs, so block b runs on XCD b%8; we hand each XCD a contiguous band
// of tiles so neighbouring halos are read through the same L2.
#include "wf_common.hpp"

namespace {

constexpr int kNT = 256;        // threads per workgroup
constexpr int kMaxGroups = 64;  // source groups per tile (host checks)
constexpr int kMaxDets = 64;    // detectors per tile core (host checks)

template <int TI, int TJ, int D, bool MAT>
__global__ void __launch_bounds__(kNT) k_tb(wf_tb_args a) {
    constexpr int RI = TI + 2 * D;
    constexpr int RJ = TJ + 2 * D;
    constexpr int NC = RI * RJ;
    __shared__ float s_ex[NC];
    __shared__ float s_ey[NC];
    __shared__ float s_hz[NC];
    __shared__ float s_ca[MAT ? NC : 1];
    __shared__ float s_cb[MAT ? NC : 1];
    __shared__ float s_prow[4][RJ];
    __shared__ float s_pcol[4][RI];
    __shared__ int s_grp[kMaxGroups];
    __shared__ int s_grp_cell[kMaxGroups];
    __shared__ int s_det[kMaxDets];
    __shared__ int s_det_cell[kMaxDets];
    __shared__ int s_ngrp, s_ndet;
    __shared__ unsigned int s_max[3];

    const int tid = threadIdx.x;
    const int per = (a.n_tiles + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= a.n_tiles) return;
    const int run = a.tiles[2 * t];
    const int tile = a.tiles[2 * t + 1];
    const wf_run2d R = a.runs[run];
    const int nx = R.nx, ny = R.ny;
    const int ntj = (ny + TJ - 1) / TJ;
    const int I0 = (tile / ntj) * TI;
    const int J0 = (tile % ntj) * TJ;
    const int gi0 = I0 - D, gj0 = J0 - D;  // global coordinates of LDS (0, 0)

    const float *src_ex = a.parity ? R.ex2 : R.ex;
    const float *src_ey = a.parity ? R.ey2 : R.ey;
    const float *src_hz = a.parity ? R.hz2 : R.hz;
    float *dst_ex = a.parity ? R.ex : R.ex2;
    float *dst_ey = a.parity ? R.ey : R.ey2;
    float *dst_hz = a.parity ? R.hz : R.hz2;

    if (tid == 0) {
        s_ngrp = 0;
        s_ndet = 0;
        s_max[0] = s_max[1] = s_max[2] = 0u;
    }
    __syncthreads();

    // Stage region and collect the sources / core detectors that touch it.
    for (int idx = tid; idx < NC; idx += kNT) {
        const int li = idx / RJ, lj = idx - (idx / RJ) * RJ;
        const int gi = gi0 + li, gj = gj0 + lj;
        float vx = 0.f, vy = 0.f, vh = 0.f, va = 0.f, vb = 0.f;
        if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
            const size_t g = (size_t)gi * ny + gj;
            vx = src_ex[g];
            vy = src_ey[g];
            vh = src_hz[g];
            if (MAT) {
                va = R.ca[g];
                vb = R.cb[g];
            }
        }
        s_ex[idx] = vx;
        s_ey[idx] = vy;
        s_hz[idx] = vh;
        if (MAT) {
            s_ca[idx] = va;
            s_cb[idx] = vb;
        }
    }
    for (int g = tid; g < R.n_grp; g += kNT) {
        const int cell = R.inj[R.grp_start[g]].cell;
        const int li = cell / ny - gi0, lj = cell % ny - gj0;
        if (li >= 0 && li < RI && lj >= 0 && lj < RJ) {
            const int k = atomicAdd(&s_ngrp, 1);
            if (k < kMaxGroups) {
                s_grp[k] = g;
                s_grp_cell[k] = li * RJ + lj;
            }
        }
    }
    for (int d = tid; d < R.n_det; d += kNT) {
        const int cell = R.det_cell[d];
        const int gi = cell / ny, gj = cell % ny;
        if (gi >= I0 && gi < I0 + TI && gj >= J0 && gj < J0 + TJ) {
            const int k = atomicAdd(&s_ndet, 1);
            if (k < kMaxDets) {
                s_det[k] = d;
                s_det_cell[k] = (gi - gi0) * RJ + (gj - gj0);
            }
        }
    }
    __syncthreads();
    const int ngrp = min(s_ngrp, kMaxGroups);
    const int ndet = min(s_ndet, kMaxDets);
    const bool edge = (gi0 <= 0) || (gi0 + RI >= nx) || (gj0 <= 0) || (gj0 + RJ >= ny);
    const float dh = R.dh, de = R.de, dx = R.dx, dy = R.dy, cx = R.cmx, cy = R.cmy;
    (void)de;

#pragma unroll
    for (int s = 1; s <= D; ++s) {
        const int n = R.step + a.step_offset + s - 1;
        // 1. Mur snapshot strips (only the edge rows/columns in this region).
        if (edge) {
            for (int q = tid; q < 4 * RJ; q += kNT) {
                const int sr = q / RJ, lj = q - sr * RJ;
                const int li = (sr < 2 ? sr : nx - 4 + sr) - gi0;
                if (li >= 0 && li < RI) s_prow[sr][lj] = s_hz[li * RJ + lj];
            }
            for (int q = tid; q < 4 * RI; q += kNT) {
                const int sc = q / RI, li = q - sc * RI;
                const int lj = (sc < 2 ? sc : ny - 4 + sc) - gj0;
                if (lj >= 0 && lj < RJ) s_pcol[sc][li] = s_hz[li * RJ + lj];
            }
            __syncthreads();
        }
        // 2. Soft sources.
        if (ngrp > 0) {
            for (int k = tid; k < ngrp; k += kNT) {
                const int g = s_grp[k];
                const int b0 = R.grp_start[g], b1 = R.grp_start[g + 1];
                const int comp = R.inj[b0].comp;
                float *f = comp == WF_COMP_EX ? s_ex : (comp == WF_COMP_EY ? s_ey : s_hz);
                const int c = s_grp_cell[k];
                float v = f[c];
                for (int e = b0; e < b1; ++e) {
                    const float w = R.inj[e].wave[n] * R.inj[e].amp;
                    v = v + w;
                }
                f[c] = v;
            }
            __syncthreads();
        }
        // 3. Faraday + Mur (owner writes).
        {
            constexpr int dummy = 0;
            (void)dummy;
            const int lo = s - 1;
            const int hi_i = RI - s, hi_j = RJ - s;
            const int w = hi_j - lo;
            const int cnt = (hi_i - lo) * w;
            for (int idx = tid; idx < cnt; idx += kNT) {
                const int li = lo + idx / w, lj = lo + (idx - (idx / w) * w);
                const int gi = gi0 + li, gj = gj0 + lj;
                if (gi < 1 || gi > nx - 2 || gj < 1 || gj > ny - 2) continue;
                const int c = li * RJ + lj;
                const float h1 = s_hz[c];
                const float av = (s_ex[c + 1] - s_ex[c]) / dy;
                const float bv = (s_ey[c + RJ] - s_ey[c]) / dx;
                const float F = h1 + dh * (av - bv);
                s_hz[c] = F;
                if (edge) {
                    if (gi == 1 && li >= 1) {
                        const float L = s_prow[1][lj] + cx * (F - s_prow[0][lj]);
                        s_hz[c - RJ] = L;
                        if (gj == 1 && lj >= 1) s_hz[c - RJ - 1] = s_prow[0][lj] + cy * (L - s_prow[0][lj - 1]);
                        if (gj == ny - 2 && lj + 1 < RJ)
                            s_hz[c - RJ + 1] = s_prow[0][lj] + cy * (L - s_prow[0][lj + 1]);
                    }
                    if (gi == nx - 2 && li + 1 < RI) {
                        const float Rv = s_prow[2][lj] + cx * (F - s_prow[3][lj]);
                        s_hz[c + RJ] = Rv;
                        if (gj == 1 && lj >= 1) s_hz[c + RJ - 1] = s_prow[3][lj] + cy * (Rv - s_prow[3][lj - 1]);
                        if (gj == ny - 2 && lj + 1 < RJ)
                            s_hz[c + RJ + 1] = s_prow[3][lj] + cy * (Rv - s_prow[3][lj + 1]);
                    }
                    if (gj == 1 && lj >= 1) s_hz[c - 1] = s_pcol[1][li] + cy * (F - s_pcol[0][li]);
                    if (gj == ny - 2 && lj + 1 < RJ) s_hz[c + 1] = s_pcol[2][li] + cy * (F - s_pcol[3][li]);
                }
            }
        }
        __syncthreads();
        // 4. Ampere + detectors.
        if (tid < ndet) {
            const int tt = n - R.trace_n0;
            if (tt >= 0 && tt < R.trace_len) R.traces[s_det[tid] * R.trace_stride + tt] = s_hz[s_det_cell[tid]];
        }
        {
            const int lo = s;
            const int hi_i = RI - s, hi_j = RJ - s;
            const int w = hi_j - lo;
            const int cnt = (hi_i - lo) * w;
            for (int idx = tid; idx < cnt; idx += kNT) {
                const int li = lo + idx / w, lj = lo + (idx - (idx / w) * w);
                const int gi = gi0 + li, gj = gj0 + lj;
                if (gi < 0 || gi >= nx || gj < 0 || gj >= ny) continue;
                const int c = li * RJ + lj;
                const float h = s_hz[c];
                if (gj >= 1) {
                    const float g = (h - s_hz[c - 1]) / dy;
                    if (MAT) {
                        const float t1 = s_ca[c] * s_ex[c];
                        const float t2 = s_cb[c] * g;
                        s_ex[c] = t1 + t2;
                    } else {
                        const float t2 = de * g;
                        s_ex[c] = s_ex[c] + t2;
                    }
                }
                if (gi >= 1) {
                    const float g = (h - s_hz[c - RJ]) / dx;
                    if (MAT) {
                        const float t1 = s_ca[c] * s_ey[c];
                        const float t2 = s_cb[c] * g;
                        s_ey[c] = t1 - t2;
                    } else {
                        const float t2 = de * g;
                        s_ey[c] = s_ey[c] - t2;
                    }
                }
            }
        }
        __syncthreads();
    }

    // Write the core back (and fold the stability maxima if requested).
    unsigned int m0 = 0u, m1 = 0u, m2 = 0u;
    for (int idx = tid; idx < TI * TJ; idx += kNT) {
        const int ci = idx / TJ, cj = idx - (idx / TJ) * TJ;
        const int gi = I0 + ci, gj = J0 + cj;
        if (gi >= nx || gj >= ny) continue;
        const int c = (ci + D) * RJ + (cj + D);
        const size_t g = (size_t)gi * ny + gj;
        const float vx = s_ex[c], vy = s_ey[c], vh = s_hz[c];
        dst_ex[g] = vx;
        dst_ey[g] = vy;
        dst_hz[g] = vh;
        m0 = max(m0, wf::abs_bits(vx));
        m1 = max(m1, wf::abs_bits(vy));
        m2 = max(m2, wf::abs_bits(vh));
    }
    if (a.maxabs_out) {
        atomicMax(&s_max[0], m0);
        atomicMax(&s_max[1], m1);
        atomicMax(&s_max[2], m2);
        __syncthreads();
        if (tid < 3) atomicMax(&a.maxabs_out[3 * run + tid], s_max[tid]);
    }
}

template <int TI, int TJ, bool MAT>
hipError_t launch_depth(const wf_tb_args &a, hipStream_t st) {
    const unsigned blocks = 8u * (unsigned)((a.n_tiles + 7) / 8);
    switch (a.depth) {
#define WF_TB_CASE(DD)                                                                      \
    case DD:                                                                                \
        hipLaunchKernelGGL((k_tb<TI, TJ, DD, MAT>), dim3(blocks), dim3(kNT), 0, st, a);     \
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

template <int TI, int TJ>
hipError_t launch_shape(const wf_tb_args &a, hipStream_t st) {
    return a.has_materials ? launch_depth<TI, TJ, true>(a, st) : launch_depth<TI, TJ, false>(a, st);
}

// (tile_i, tile_j) pairs compiled into the library.
constexpr int kShapes[][2] = {{32, 64}, {16, 64}, {64, 64}, {32, 32}};

}  // namespace

extern "C" int wf_tb_tile_shapes(int32_t *out, int32_t cap) {
    const int n = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
    for (int k = 0; k < n && k < cap; ++k) {
        out[2 * k] = kShapes[k][0];
        out[2 * k + 1] = kShapes[k][1];
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
    if (a.parity != 0 && a.parity != 1) {
        wf::set_error("wf_fdtd2d_tb_advance: parity must be 0 or 1");
        return WF_ERR_ARG;
    }
    hipError_t e;
    if (a.tile_i == 32 && a.tile_j == 64)
        e = launch_shape<32, 64>(a, stream);
    else if (a.tile_i == 16 && a.tile_j == 64)
        e = launch_shape<16, 64>(a, stream);
    else if (a.tile_i == 64 && a.tile_j == 64)
        e = launch_shape<64, 64>(a, stream);
    else if (a.tile_i == 32 && a.tile_j == 32)
        e = launch_shape<32, 32>(a, stream);
    else {
        wf::set_error("wf_fdtd2d_tb_advance: unsupported tile shape %dx%d", a.tile_i, a.tile_j);
        return WF_ERR_UNSUPPORTED;
    }
    return wf::check_hip(e, "wf_fdtd2d_tb_advance");
}
