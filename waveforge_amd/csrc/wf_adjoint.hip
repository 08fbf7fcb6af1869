// Discrete adjoint of one 2D TM leapfrog step, for gradients of a loss on the
// recorded Hz traces with respect to the per-cell coefficients Ca, Cb.
//
// The forward step n (src/core/fdtd2d.py:187-266) maps the state u_n =
// (Ex, Ey, Hz) to u_{n+1} through, in order: snapshot P = Hz (S1), soft
// sources (S2, affine), Faraday on [0, nx-1) x [0, ny-1) (S3), Mur left,
// right, bottom, top (S4-S7, src/core/boundaries.py:259-274), Ampere for Ex
// and Ey (S8, S9), and the record d_n = Hz_{n+1}[rx]. Every stage is linear in
// the state, so the adjoint applies the transposed stages in reverse order to
// the adjoint state lam = (lEx, lEy, lHz) of u_{n+1} and yields the adjoint of
// u_n, accumulating dL/dCa = sum_n lE_{n+1} * E_n and dL/dCb = sum_n lE_{n+1} * curl_n:
//
//   k_adj_record  lHz[rx] += dL/dd_n[rx] (with a receiver map, only for the
//                 first step of a call; later steps add theirs in the step
//                 kernel before)
//   k_adj_eh      the whole step in one launch (the default): the two stages
//                 below over a tile and a ring around it, through LDS
//   k_adj_e       transposed Ampere (S9, S8) + coefficient gradients;
//                 lHz gathers the four Ampere neighbours (out of place)
//   k_adj_h       transposed Mur (S7..S4) and Faraday (S3) gathered per
//                 cell, plus the snapshot's transpose lHz += lP (S1)
//
// All runs of a batch share the grid size and Ca/Cb (one phantom, many TX):
// blockIdx.z is the run, and run r's arrays start at r * run_stride. Gradients
// accumulate per run in fp64. This is a correct adjoint of the fp32 forward's
// exact-arithmetic counterpart; it is checked against autograd through the
// oracle and against central differences.
#include "wf_common.hpp"
#include "wf_div.hpp"

namespace {

constexpr int kBX = 64;
constexpr int kBY = 4;

__device__ __forceinline__ float residual(const wf_adj_args &a, int r, int k, int step) {
    return a.dl_dd[(int64_t)r * a.dd_run_stride + (int64_t)k * a.dd_rx_stride + step];
}

// lHz at one cell plus the residuals of the receivers there, in receiver order
// (rx_order lists them by cell, rx_first[c] is the first position of cell c).
__device__ __forceinline__ float add_residuals(const wf_adj_args &a, int r, int64_t c, int p, int step, float v) {
    for (; p < a.n_rx && a.rx_cell[a.rx_order[p]] == c; ++p) v = v + residual(a, r, a.rx_order[p], step);
    return v;
}

__global__ void k_adj_record(wf_adj_args a) {
    const int r = blockIdx.x;
    const int64_t off = (int64_t)r * a.run_stride;
    if (a.rx_first) {  // one thread per receiver cell, its receivers in order
        for (int p = threadIdx.x; p < a.n_rx; p += blockDim.x) {
            const int c = a.rx_cell[a.rx_order[p]];
            if (p > 0 && a.rx_cell[a.rx_order[p - 1]] == c) continue;
            a.lhz[off + c] = add_residuals(a, r, c, p, a.step, a.lhz[off + c]);
        }
        return;
    }
    for (int k = threadIdx.x; k < a.n_rx; k += blockDim.x)
        atomicAdd(&a.lhz[off + a.rx_cell[k]], residual(a, r, k, a.step));  // receivers may share a cell
}

__global__ void __launch_bounds__(kBX * kBY) k_adj_e(wf_adj_args a) {
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    const int nx = a.nx, ny = a.ny;
    if (i >= nx || j >= ny) return;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride;
    const int64_t soff = (int64_t)blockIdx.z * a.state_stride;
    const int64_t c = (int64_t)i * ny + j;
    const float lex = a.lex[off + c];
    const float ley = a.ley[off + c];
    const float ca = a.ca[c], cb = a.cb[c];
    float lhz = a.lhz[off + c];
    double gca = 0.0, gcb = 0.0;
    float lex_o = lex, ley_o = ley;
    const float *hz1 = a.hz_next + soff;
    if (j >= 1) {  // S8: Ex' = Ca Ex + Cb (Hz1[i,j] - Hz1[i,j-1]) / dy
        const float gy = (hz1[c] - hz1[c - 1]) / a.dy;
        gca += (double)lex * (double)a.ex_prev[soff + c];
        gcb += (double)lex * (double)gy;
        lex_o = ca * lex;
        lhz += cb * lex / a.dy;
    }
    if (j + 1 < ny) lhz -= a.cb[c + 1] * a.lex[off + c + 1] / a.dy;
    if (i >= 1) {  // S9: Ey' = Ca Ey - Cb (Hz1[i,j] - Hz1[i-1,j]) / dx
        const float gx = (hz1[c] - hz1[c - ny]) / a.dx;
        gca += (double)ley * (double)a.ey_prev[soff + c];
        gcb -= (double)ley * (double)gx;
        ley_o = ca * ley;
        lhz -= cb * ley / a.dx;
    }
    if (i + 1 < nx) lhz += a.cb[c + ny] * a.ley[off + c + ny] / a.dx;
    a.lex2[off + c] = lex_o;
    a.ley2[off + c] = ley_o;
    a.lhz2[off + c] = lhz;
    a.gca[off + c] += gca;
    a.gcb[off + c] += gcb;
}

// Transposed Faraday (S3) with the transposed Mur (S7..S4) folded in: the Mur
// transpose only rewrites the two outer rows and columns of lHz' (the output
// of the transposed Ampere) and accumulates the snapshot's adjoint lP there,
// with lP zero beforehand, so each cell's post-Mur lHz' and lP are short
// expressions in at most four lHz' values, formed here in the same fp32
// operation order as the sequential transpose: top then bottom face of every
// row (S7, S6), then right then left face of every column (S5, S4). Returns
// F(i, j), F(i, j-1), F(i-1, j) with F the post-Mur lHz' on Faraday's range
// [0, nx-1) x [0, ny-1), the post-Mur lHz'(i, j) and lP(i, j); H(ii, jj)
// reads lHz' and is called only within two cells of (i, j).
template <class HF>
__device__ __forceinline__ void mur_faraday_t(const HF &H, int i, int j, int nx, int ny, float cx, float cy,
                                              float &fij, float &fl, float &fd, float &hz, float &pv) {
    // after the row faces (S7 then S6 transposed)
    auto H1 = [&](int ii, int jj) -> float {
        if (jj == ny - 1 || jj == 0) return 0.f;
        if (jj == ny - 2) return H(ii, ny - 2) + cy * H(ii, ny - 1);
        if (jj == 1) return H(ii, 1) + cy * H(ii, 0);
        return H(ii, jj);
    };
    // after the column faces (S5 then S4 transposed)
    auto H2 = [&](int ii, int jj) -> float {
        if (ii == nx - 1 || ii == 0) return 0.f;
        if (ii == nx - 2) return H1(nx - 2, jj) + cx * H1(nx - 1, jj);
        if (ii == 1) return H1(1, jj) + cx * H1(0, jj);
        return H1(ii, jj);
    };
    auto F = [&](int ii, int jj) -> float {
        return (ii >= 0 && jj >= 0 && ii <= nx - 2 && jj <= ny - 2) ? H2(ii, jj) : 0.f;
    };
    pv = 0.f;
    const bool mid_rows = i >= 3 && i <= nx - 3;  // (i, j), (i-1, j) on no column face
    if (mid_rows && j >= 3 && j <= ny - 3) {      // no face term reaches (i, j), (i, j-1), (i-1, j)
        fij = H(i, j);
        fl = H(i, j - 1);
        fd = H(i - 1, j);
        hz = fij;
    } else if (mid_rows) {  // row faces only: cheap, for the lanes at the ends of every row
        auto F1 = [&](int ii, int jj) -> float { return (jj >= 0 && jj <= ny - 2) ? H1(ii, jj) : 0.f; };
        fij = F1(i, j);
        fl = F1(i, j - 1);
        fd = F1(i - 1, j);
        hz = H1(i, j);
        if (j == ny - 2) pv = pv + H(i, ny - 1);
        if (j == ny - 1) pv = pv - cy * H(i, ny - 1);
        if (j == 1) pv = pv + H(i, 0);
        if (j == 0) pv = pv - cy * H(i, 0);
    } else {
        fij = F(i, j);
        fl = F(i, j - 1);
        fd = F(i - 1, j);
        hz = H2(i, j);
        if (j == ny - 2) pv = pv + H(i, ny - 1);
        if (j == ny - 1) pv = pv - cy * H(i, ny - 1);
        if (j == 1) pv = pv + H(i, 0);
        if (j == 0) pv = pv - cy * H(i, 0);
        if (i == nx - 2) pv = pv + H1(nx - 1, j);
        if (i == nx - 1) pv = pv - cx * H1(nx - 1, j);
        if (i == 1) pv = pv + H1(0, j);
        if (i == 0) pv = pv - cx * H1(0, j);
    }
}

// next_step >= 0: also record the receivers of adjoint step next_step (the
// one after this in the backward sweep) into the lHz written here.
__global__ void __launch_bounds__(kBX * kBY) k_adj_h(wf_adj_args a, int next_step) {
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    const int nx = a.nx, ny = a.ny;
    if (i >= nx || j >= ny) return;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride;
    const int64_t c = (int64_t)i * ny + j;
    const float *h = a.lhz2 + off;
    auto H = [&](int ii, int jj) -> float { return h[(int64_t)ii * ny + jj]; };
    float fij, fl, fd, hz, pv;
    mur_faraday_t(H, i, j, nx, ny, a.cmx, a.cmy, fij, fl, fd, hz, pv);
    a.lex[off + c] = a.lex2[off + c] + (a.dh / a.dy) * (fl - fij);
    a.ley[off + c] = a.ley2[off + c] + (a.dh / a.dx) * (fij - fd);
    float out = hz + pv;
    if (next_step >= 0) {
        const int p = a.rx_first[c];
        if (p >= 0) out = add_residuals(a, blockIdx.z, c, p, next_step, out);
    }
    a.lhz[off + c] = out;
}

// One whole adjoint step in one launch: the transposed Ampere (k_adj_e's
// arithmetic, in its order) over the tile and a two-cell ring around it into
// LDS, then the transposed Mur and Faraday (k_adj_h's) over the tile from LDS.
// Reads the adjoint state from st.lex/ley/lhz and writes it to
// st.olex/oley/olhz (the ring reads cells that neighbouring tiles rewrite, so
// the step cannot be in place); the coefficient gradients accumulate as in
// k_adj_e. Saves k_adj_e's three intermediate fields: their write and
// k_adj_h's read of them. Wavefront w holds region rows w, w + kEhWaves, ...
// (one 64-lane row each).
constexpr int kEhRows = 8;                // tile rows
constexpr int kEhCols = kBX - 4;          // tile columns (lanes 2..61 of the 64-lane region)
constexpr int kEhRegion = kEhRows + 4;    // region rows
constexpr int kEhWaves = 4;               // wavefronts per workgroup
constexpr int kEhPerWave = (kEhRegion + kEhWaves - 1) / kEhWaves;

struct AdjState {
    const float *lex, *ley, *lhz;
    float *olex, *oley, *olhz;
};

__global__ void __launch_bounds__(kBX * kEhWaves) k_adj_eh(wf_adj_args a, AdjState st, int next_step) {
    __shared__ float s_h[kEhRegion][kBX];
    const int q = threadIdx.x, w = threadIdx.y;
    const int nx = a.nx, ny = a.ny;
    const int i0 = blockIdx.y * kEhRows - 2, j0 = blockIdx.x * kEhCols - 2;  // region origin
    const int j = j0 + q;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride;
    const int64_t soff = (int64_t)blockIdx.z * a.state_stride;
    const float *hz1 = a.hz_next + soff;
    const bool core_col = q >= 2 && q < 2 + kEhCols && j < ny;
    // The transposed Ampere's neighbour terms are the neighbours' own terms:
    // vx = (Cb lEx)/dy of cell (i, j+1) and vy = (Cb lEy)/dx of (i+1, j).
    // Each cell forms its vx, vy once; vx reaches the cell to its left by a
    // lane shift, vy the cell above through LDS. Loads come first (the
    // gradient stores below could alias them as far as the compiler knows).
    float lex[kEhPerWave], ley[kEhPerWave], lhz[kEhPerWave], ca[kEhPerWave];
    float vx[kEhPerWave], vy[kEhPerWave];
#pragma unroll
    for (int t = 0; t < kEhPerWave; ++t) {
        const int p = w + kEhWaves * t;
        const int i = i0 + p;
        lex[t] = ley[t] = lhz[t] = ca[t] = 0.f;
        float cb = 0.f;
        if (p < kEhRegion && i >= 0 && i < nx && j >= 0 && j < ny) {
            const int64_t c = (int64_t)i * ny + j;
            lex[t] = st.lex[off + c];
            ley[t] = st.ley[off + c];
            lhz[t] = st.lhz[off + c];
            ca[t] = a.ca[c];
            cb = a.cb[c];
        }
        vx[t] = cb * lex[t] / a.dy;
        vy[t] = cb * ley[t] / a.dx;
        if (p < kEhRegion) s_h[p][q] = vy[t];
    }
    // the inputs of the tile cells' coefficient gradients, in flight with the
    // loads above (the sums and stores come after the stages below)
    float h0[kEhPerWave], hl[kEhPerWave], hu[kEhPerWave], exp_[kEhPerWave], eyp[kEhPerWave];
    double g0[kEhPerWave], g1[kEhPerWave];
#pragma unroll
    for (int t = 0; t < kEhPerWave; ++t) {
        const int p = w + kEhWaves * t;
        const int i = i0 + p;
        h0[t] = hl[t] = hu[t] = exp_[t] = eyp[t] = 0.f;
        g0[t] = g1[t] = 0.0;
        if (core_col && p >= 2 && p < 2 + kEhRows && i < nx) {
            const int64_t c = (int64_t)i * ny + j;
            h0[t] = hz1[c];
            if (j >= 1) {
                hl[t] = hz1[c - 1];
                exp_[t] = a.ex_prev[soff + c];
            }
            if (i >= 1) {
                hu[t] = hz1[c - ny];
                eyp[t] = a.ey_prev[soff + c];
            }
            g0[t] = a.gca[off + c];
            g1[t] = a.gcb[off + c];
        }
    }
    __syncthreads();
    float lz[kEhPerWave];
#pragma unroll
    for (int t = 0; t < kEhPerWave; ++t) {
        const int p = w + kEhWaves * t;
        lz[t] = lhz[t];
        const float vx_r = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(vx[t]), 0x130, 0xf, 0xf, false));
        const float vy_d = p + 1 < kEhRegion ? s_h[p + 1][q] : 0.f;
        const int i = i0 + p;
        if (j >= 1) lz[t] += vx[t];
        if (j + 1 < ny) lz[t] -= vx_r;
        if (i >= 1) lz[t] -= vy[t];
        if (i + 1 < nx) lz[t] += vy_d;
    }
    __syncthreads();
    float lex_o[kEhPerWave], ley_o[kEhPerWave];
#pragma unroll
    for (int t = 0; t < kEhPerWave; ++t) {
        const int p = w + kEhWaves * t;
        const int i = i0 + p;
        lex_o[t] = ley_o[t] = 0.f;
        if (p >= kEhRegion) continue;
        float out = 0.f;
        if (i >= 0 && i < nx && j >= 0 && j < ny) {
            const int64_t c = (int64_t)i * ny + j;
            out = lz[t];
            if (core_col && p >= 2 && p < 2 + kEhRows) {
                double gca = 0.0, gcb = 0.0;
                lex_o[t] = lex[t];
                ley_o[t] = ley[t];
                if (j >= 1) {  // S8: Ex' = Ca Ex + Cb (Hz1[i,j] - Hz1[i,j-1]) / dy
                    const float gy = (h0[t] - hl[t]) / a.dy;
                    gca += (double)lex[t] * (double)exp_[t];
                    gcb += (double)lex[t] * (double)gy;
                    lex_o[t] = ca[t] * lex[t];
                }
                if (i >= 1) {  // S9: Ey' = Ca Ey - Cb (Hz1[i,j] - Hz1[i-1,j]) / dx
                    const float gx = (h0[t] - hu[t]) / a.dx;
                    gca += (double)ley[t] * (double)eyp[t];
                    gcb -= (double)ley[t] * (double)gx;
                    ley_o[t] = ca[t] * ley[t];
                }
                a.gca[off + c] = g0[t] + gca;
                a.gcb[off + c] = g1[t] + gcb;
            }
        }
        s_h[p][q] = out;
    }
    __syncthreads();
    auto H = [&](int ii, int jj) -> float { return s_h[ii - i0][jj - j0]; };
#pragma unroll
    for (int t = 0; t < kEhPerWave; ++t) {
        const int p = w + kEhWaves * t;
        const int i = i0 + p;
        if (p < 2 || p >= 2 + kEhRows || i >= nx || !core_col) continue;
        const int64_t c = (int64_t)i * ny + j;
        float fij, fl, fd, hz, pv;
        mur_faraday_t(H, i, j, nx, ny, a.cmx, a.cmy, fij, fl, fd, hz, pv);
        st.olex[off + c] = lex_o[t] + (a.dh / a.dy) * (fl - fij);
        st.oley[off + c] = ley_o[t] + (a.dh / a.dx) * (fij - fd);
        float o = hz + pv;
        if (next_step >= 0) {
            const int r0 = a.rx_first[c];
            if (r0 >= 0) o = add_residuals(a, blockIdx.z, c, r0, next_step, o);
        }
        st.olhz[off + c] = o;
    }
}

// Two adjoint steps (s, then s-1) in one launch. One step's output at a cell
// reads its input within two cells (the transposed Ampere reaches +1 in i and
// j, the transposed Mur/Faraday -2..+1), so a tile of kE2Rows x kE2Cols cells
// loads a region with a four-cell ring; step A forms the state of step s-1
// over the tile and a two-cell ring (the ring cells as the neighbouring tiles
// form them: the same operations on the same inputs), keeps it in registers,
// and step B forms the state of step s-2 over the tile from it. Every cell
// value is formed by k_adj_eh's operations in its order, and a tile cell's
// gradient is (g + g_A) + g_B as two k_adj_eh launches leave it, so the bits
// equal two one-step launches. The adjoint state is read and written once per
// two steps, Ca/Cb read once, the fp64 gradients read and written once.
// mid_step >= 0: the record of step s-1 (receiver map required) is added to
// step A's lHz; next_step >= 0: that of step s-2 to the output lHz.
// A value at a byte offset from a base pointer (a uniform base and a 32-bit
// per-lane offset: the loads and stores take the scalar-base address form).
template <class T>
__device__ __forceinline__ T ld(const T *base, unsigned int bo) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + bo);
}
template <class T>
__device__ __forceinline__ void st_(T *base, unsigned int bo, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + bo) = v;
}

// FAST: the divisions by dx and dy use the verified fast quotient (the same
// bits as IEEE division). ROWS: tile rows (region ROWS + 8 over 8 wavefronts).
constexpr int kE2Cols = kBX - 8;          // tile columns (lanes 4..59)
constexpr int kE2Waves = 8;               // wavefronts per workgroup

template <bool FAST, int ROWS>
__global__ void __launch_bounds__(kBX * kE2Waves) k_adj_eh2(wf_adj_args a, AdjState st, int mid_step, int next_step) {
    constexpr int kE2Rows = ROWS;
    constexpr int kE2Region = kE2Rows + 8;
    constexpr int kE2PerWave = kE2Region / kE2Waves;
    static_assert(kE2Region % kE2Waves == 0, "region rows split evenly over the wavefronts");
    __shared__ float s_h[kE2Region][kBX];
    wf::Divider divx{}, divy{};
    divx.d = a.dx;  // the IEEE form divides by d
    divy.d = a.dy;
    if (FAST) {
        divx = wf::make_divider(a.dx, a.div_rx, a.div_span_x);
        divy = wf::make_divider(a.dy, a.div_ry, a.div_span_y);
    }
    const int q = threadIdx.x, w = threadIdx.y;
    const int nx = a.nx, ny = a.ny;
    const int i0 = blockIdx.y * kE2Rows - 4, j0 = blockIdx.x * kE2Cols - 4;  // region origin
    const int j = j0 + q;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride;
    const int64_t soff = (int64_t)blockIdx.z * a.state_stride;
    const bool jin = j >= 0 && j < ny;
    const bool a_col = q >= 2 && q < kBX - 2 && jin;            // step A's output columns
    const bool core_col = q >= 4 && q < 4 + kE2Cols && j < ny;  // the tile's columns (j >= 0 there)
    const float *hzA = a.hz_next + soff, *exA = a.ex_prev + soff, *eyA = a.ey_prev + soff;
    const float *hzB = hzA - a.hist_stride, *exB = exA - a.hist_stride, *eyB = eyA - a.hist_stride;
    auto in_grid = [&](int i) { return i >= 0 && i < nx && jin; };
    auto core_row = [&](int p, int i) { return p >= 4 && p < 4 + kE2Rows && i < nx; };

    // every global load first (the stores at the end could alias them as far
    // as the compiler knows): the adjoint state and coefficients of the
    // region, the receiver map of step A's cells, and both steps' gradient
    // inputs and accumulators of the tile. Arrays are addressed as a uniform
    // base plus a 32-bit byte offset (the host keeps nx*ny*8 below 2^32), one
    // offset per row for all of them.
    const float *b_lex = st.lex + off, *b_ley = st.ley + off, *b_lhz = st.lhz + off;
    const double *b_gca = a.gca + off, *b_gcb = a.gcb + off;
    float lex[kE2PerWave], ley[kE2PerWave], lhz[kE2PerWave], ca[kE2PerWave], cb[kE2PerWave];
    float h0A[kE2PerWave], hlA[kE2PerWave], huA[kE2PerWave], exA_[kE2PerWave], eyA_[kE2PerWave];
    float h0B[kE2PerWave], hlB[kE2PerWave], huB[kE2PerWave], exB_[kE2PerWave], eyB_[kE2PerWave];
    double g0[kE2PerWave], g1[kE2PerWave];
    int rxf[kE2PerWave];
    const bool recs = mid_step >= 0 || next_step >= 0;
#pragma unroll
    for (int t = 0; t < kE2PerWave; ++t) {
        const int p = w + kE2Waves * t;
        const int i = i0 + p;
        const unsigned int c4 = (unsigned int)(i * ny + j) * 4u;  // byte offset of a float (unused outside the grid)
        lex[t] = ley[t] = lhz[t] = ca[t] = cb[t] = 0.f;
        if (in_grid(i)) {
            lex[t] = ld(b_lex, c4);
            ley[t] = ld(b_ley, c4);
            lhz[t] = ld(b_lhz, c4);
            ca[t] = ld(a.ca, c4);
            cb[t] = ld(a.cb, c4);
        }
        rxf[t] = -1;
        if (recs && p >= 2 && p < kE2Region - 2 && a_col && i >= 0 && i < nx) rxf[t] = ld(a.rx_first, c4);
        h0A[t] = hlA[t] = huA[t] = exA_[t] = eyA_[t] = 0.f;
        h0B[t] = hlB[t] = huB[t] = exB_[t] = eyB_[t] = 0.f;
        g0[t] = g1[t] = 0.0;
        if (core_col && core_row(p, i)) {
            h0A[t] = ld(hzA, c4);
            h0B[t] = ld(hzB, c4);
            if (j >= 1) {
                hlA[t] = ld(hzA, c4 - 4u);
                exA_[t] = ld(exA, c4);
                hlB[t] = ld(hzB, c4 - 4u);
                exB_[t] = ld(exB, c4);
            }
            if (i >= 1) {
                huA[t] = ld(hzA, c4 - 4u * (unsigned int)ny);
                eyA_[t] = ld(eyA, c4);
                huB[t] = ld(hzB, c4 - 4u * (unsigned int)ny);
                eyB_[t] = ld(eyB, c4);
            }
            g0[t] = ld(b_gca, 2u * c4);
            g1[t] = ld(b_gcb, 2u * c4);
        }
    }
    auto H = [&](int ii, int jj) -> float { return s_h[ii - i0][jj - j0]; };
    float lex_o[kE2PerWave], ley_o[kE2PerWave];

    // one step over rows [lo, hi) of the region and lanes [2, 62) (or the
    // tile's lanes for the last step): the transposed Ampere into LDS, the
    // gradient terms of the tile cells (h*, ex*, ey* of that step; added to
    // g0, g1), then the transposed Mur/Faraday; the result replaces lex, ley,
    // lhz (zero outside the rows and columns formed).
    auto adj_step = [&](const float *h0, const float *hl, const float *hu, const float *ex_, const float *ey_,
                        int lo, int hi, bool last, int rec_step) {
        float vx[kE2PerWave], vy[kE2PerWave], px[kE2PerWave], py[kE2PerWave];
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) {
            px[t] = cb[t] * lex[t];
            py[t] = cb[t] * ley[t];
        }
        wf::divide_row<(FAST ? wf::kDivFast : wf::kDivIeee), kE2PerWave>(px, divy, py, divx, vx, vy);
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) s_h[w + kE2Waves * t][q] = vy[t];
        __syncthreads();
        float lz[kE2PerWave];
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) {
            const int p = w + kE2Waves * t;
            lz[t] = lhz[t];
            const float vx_r =
                __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(vx[t]), 0x130, 0xf, 0xf, false));
            const float vy_d = p + 1 < kE2Region ? s_h[p + 1][q] : 0.f;
            const int i = i0 + p;
            if (j >= 1) lz[t] += vx[t];
            if (j + 1 < ny) lz[t] -= vx_r;
            if (i >= 1) lz[t] -= vy[t];
            if (i + 1 < nx) lz[t] += vy_d;
        }
        __syncthreads();
        // the gradient's differences over dy and dx (used at the tile cells)
        float dgy[kE2PerWave], dgx[kE2PerWave], gyq[kE2PerWave], gxq[kE2PerWave];
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) {
            dgy[t] = h0[t] - hl[t];
            dgx[t] = h0[t] - hu[t];
        }
        wf::divide_row<(FAST ? wf::kDivFast : wf::kDivIeee), kE2PerWave>(dgy, divy, dgx, divx, gyq, gxq);
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) {
            const int p = w + kE2Waves * t;
            const int i = i0 + p;
            lex_o[t] = lex[t];
            ley_o[t] = ley[t];
            float out = 0.f;
            if (in_grid(i)) {
                out = lz[t];
                double gca = 0.0, gcb = 0.0;
                const bool tile = core_col && core_row(p, i);
                if (j >= 1) {  // S8: Ex' = Ca Ex + Cb (Hz1[i,j] - Hz1[i,j-1]) / dy
                    if (tile) {
                        const float gy = gyq[t];
                        gca += (double)lex[t] * (double)ex_[t];
                        gcb += (double)lex[t] * (double)gy;
                    }
                    lex_o[t] = ca[t] * lex[t];
                }
                if (i >= 1) {  // S9: Ey' = Ca Ey - Cb (Hz1[i,j] - Hz1[i-1,j]) / dx
                    if (tile) {
                        const float gx = gxq[t];
                        gca += (double)ley[t] * (double)ey_[t];
                        gcb -= (double)ley[t] * (double)gx;
                    }
                    ley_o[t] = ca[t] * ley[t];
                }
                if (tile) {
                    g0[t] = g0[t] + gca;
                    g1[t] = g1[t] + gcb;
                }
            }
            s_h[p][q] = out;
        }
        __syncthreads();
        const bool cols = last ? core_col : a_col;
#pragma unroll
        for (int t = 0; t < kE2PerWave; ++t) {
            const int p = w + kE2Waves * t;
            const int i = i0 + p;
            float nlex = 0.f, nley = 0.f, nlhz = 0.f;
            if (p >= lo && p < hi && cols && i >= 0 && i < nx) {
                const int64_t c = (int64_t)i * ny + j;
                float fij, fl, fd, hz, pv;
                mur_faraday_t(H, i, j, nx, ny, a.cmx, a.cmy, fij, fl, fd, hz, pv);
                nlex = lex_o[t] + (a.dh / a.dy) * (fl - fij);
                nley = ley_o[t] + (a.dh / a.dx) * (fij - fd);
                nlhz = hz + pv;
                if (rec_step >= 0 && rxf[t] >= 0) nlhz = add_residuals(a, blockIdx.z, c, rxf[t], rec_step, nlhz);
            }
            lex[t] = nlex;
            ley[t] = nley;
            lhz[t] = nlhz;
        }
        __syncthreads();  // s_h is rewritten by the next step
    };

    adj_step(h0A, hlA, huA, exA_, eyA_, 2, kE2Region - 2, false, mid_step);
    adj_step(h0B, hlB, huB, exB_, eyB_, 4, 4 + kE2Rows, true, next_step);
#pragma unroll
    for (int t = 0; t < kE2PerWave; ++t) {
        const int p = w + kE2Waves * t;
        const int i = i0 + p;
        if (!(core_col && core_row(p, i))) continue;
        const unsigned int c4 = (unsigned int)(i * ny + j) * 4u;
        st_(st.olex + off, c4, lex[t]);
        st_(st.oley + off, c4, ley[t]);
        st_(st.olhz + off, c4, lhz[t]);
        st_(a.gca + off, 2u * c4, g0[t]);
        st_(a.gcb + off, 2u * c4, g1[t]);
    }
}

}  // namespace

extern "C" int wf_adjoint_step(const wf_adj_args *args, hipStream_t stream) {
    if (!args || !args->ca || !args->cb || !args->lex || !args->ley || !args->lhz || !args->lex2 || !args->ley2 ||
        !args->lhz2 || !args->lp || !args->gca || !args->gcb || !args->ex_prev || !args->ey_prev || !args->hz_next) {
        wf::set_error("wf_adjoint_step: null pointer");
        return WF_ERR_ARG;
    }
    const wf_adj_args &a = *args;
    if (a.nx < 4 || a.ny < 4 || a.n_runs < 1 || a.n_runs > 65535 || a.n_rx < 0 ||
        (a.n_rx > 0 && (!a.rx_cell || !a.dl_dd)) || (a.rx_first != nullptr) != (a.rx_order != nullptr)) {
        wf::set_error("wf_adjoint_step: bad sizes");
        return WF_ERR_ARG;
    }
    if ((a.flags & WF_ADJ_FAST_DIV) &&
        (!(a.div_rx > 0.f) || !(a.div_ry > 0.f) || !wf::divider_ok(a.dx) || !wf::divider_ok(a.dy))) {
        wf::set_error("wf_adjoint_step: WF_ADJ_FAST_DIV needs dx, dy normal with exponent in [-95, 31] and their "
                      "verified reciprocals");
        return WF_ERR_ARG;
    }
    const int n_steps = a.n_steps > 1 ? a.n_steps : 1;
    if (n_steps > 1 && (a.hist_stride <= 0 || a.step - (n_steps - 1) < 0)) {
        wf::set_error("wf_adjoint_step: %d steps need a positive hist_stride and step >= n_steps-1", n_steps);
        return WF_ERR_ARG;
    }
    wf_adj_args b = a;
    if (!(a.flags & WF_ADJ_TWO_PASS)) {
        // one launch per step (k_adj_eh) or per two steps (k_adj_eh2), the
        // adjoint state alternating between the primary fields (lex, ley,
        // lhz) and the second set (lex2, ley2, lhz2) at every launch. Two
        // steps share a launch when the receivers' records of the steps after
        // the first can be folded in (no receivers, or the receiver map).
        const dim3 ge((a.ny + kEhCols - 1) / kEhCols, (a.nx + kEhRows - 1) / kEhRows, a.n_runs);
        // 8-row tiles (a 16-row region, two rows per wavefront, 68 VGPRs, 7
        // waves per SIMD): the step is latency-bound, and the extra ring work
        // of the smaller tile costs less than the residency it buys (73 ms
        // against 79 ms for the 256^2 x 16-run, 3000-step gradient on 16-row
        // tiles at 5 waves per SIMD)
        constexpr int kRows2 = 8;
        const dim3 ge2((a.ny + kE2Cols - 1) / kE2Cols, (a.nx + kRows2 - 1) / kRows2, a.n_runs);
        const bool fast = (a.flags & WF_ADJ_FAST_DIV) != 0;
        const AdjState pri{a.lex, a.ley, a.lhz, a.lex2, a.ley2, a.lhz2};
        const AdjState sec{a.lex2, a.ley2, a.lhz2, a.lex, a.ley, a.lhz};
        const bool fold = a.n_rx > 0 && a.rx_first != nullptr;
        // two-step launches address a run's arrays by 32-bit byte offsets (fp64: 8 bytes a cell)
        const bool fits = (int64_t)a.nx * a.ny * 8 < ((int64_t)1 << 32);
        const bool pairs = !(a.flags & WF_ADJ_ONE_STEP) && (a.n_rx == 0 || fold) && fits;
        int launches = 0;
        for (int t = 0; t < n_steps; ++launches) {
            const int64_t back = (int64_t)t * a.hist_stride;
            b.step = a.step - t;
            b.ex_prev = a.ex_prev - back;
            b.ey_prev = a.ey_prev - back;
            b.hz_next = a.hz_next - back;
            if (b.n_rx > 0 && (t == 0 || !fold)) {
                wf_adj_args r = b;
                if (launches & 1) r.lhz = a.lhz2;  // the state this step reads
                hipLaunchKernelGGL(k_adj_record, dim3(b.n_runs), dim3(64), 0, stream, r);
            }
            const AdjState &sd = (launches & 1) ? sec : pri;
            if (pairs && t + 1 < n_steps) {
                const int mid = fold ? b.step - 1 : -1;
                const int next = fold && t + 2 < n_steps ? b.step - 2 : -1;
                if (fast)
                    hipLaunchKernelGGL((k_adj_eh2<true, kRows2>), ge2, dim3(kBX, kE2Waves), 0, stream, b, sd, mid, next);
                else
                    hipLaunchKernelGGL((k_adj_eh2<false, kRows2>), ge2, dim3(kBX, kE2Waves), 0, stream, b, sd, mid, next);
                t += 2;
            } else {
                const int next = fold && t + 1 < n_steps ? b.step - 1 : -1;
                hipLaunchKernelGGL(k_adj_eh, ge, dim3(kBX, kEhWaves), 0, stream, b, sd, next);
                t += 1;
            }
        }
        const int rc = wf::check_launch("wf_adjoint_step");
        if (rc || !(launches & 1)) return rc;
        // the state ended in the second set
        const size_t bytes = (size_t)a.n_runs * (size_t)a.run_stride * sizeof(float);
        hipError_t e = hipMemcpyAsync(a.lex, a.lex2, bytes, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(a.ley, a.ley2, bytes, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(a.lhz, a.lhz2, bytes, hipMemcpyDeviceToDevice, stream);
        return wf::check_hip(e, "wf_adjoint_step");
    }
    const dim3 g((a.ny + kBX - 1) / kBX, (a.nx + kBY - 1) / kBY, a.n_runs);
    for (int t = 0; t < n_steps; ++t) {
        const int64_t back = (int64_t)t * a.hist_stride;
        b.step = a.step - t;
        b.ex_prev = a.ex_prev - back;
        b.ey_prev = a.ey_prev - back;
        b.hz_next = a.hz_next - back;
        // with the receiver map, only the call's first step records on its
        // own; every later step's record rides in the k_adj_h before it
        const bool fold = b.n_rx > 0 && b.rx_first != nullptr;
        if (b.n_rx > 0 && (t == 0 || !fold)) hipLaunchKernelGGL(k_adj_record, dim3(b.n_runs), dim3(64), 0, stream, b);
        hipLaunchKernelGGL(k_adj_e, g, dim3(kBX, kBY), 0, stream, b);
        const int next = fold && t + 1 < n_steps ? b.step - 1 : -1;
        hipLaunchKernelGGL(k_adj_h, g, dim3(kBX, kBY), 0, stream, b, next);
    }
    return wf::check_launch("wf_adjoint_step");
}
