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
//                 first step of a call; later steps add theirs in k_adj_h)
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
// of k_adj_e) and accumulates the snapshot's adjoint lP there, with lP zero
// beforehand, so each cell's post-Mur lHz' and lP are short expressions in at
// most four lHz' values, formed here in the same fp32 operation order as the
// sequential transpose: top then bottom face of every row (S7, S6), then right
// then left face of every column (S5, S4). Then
//   lEx  = lEx' + (Dh/dy) (F(i, j-1) - F(i, j)),  lEy = lEy' + (Dh/dx) (F(i, j) - F(i-1, j))
// with F the post-Mur lHz' on Faraday's range [0, nx-1) x [0, ny-1), and
// lHz = post-Mur lHz' + lP (the snapshot's transpose, S1).
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
    const float cx = a.cmx, cy = a.cmy;
    auto H = [&](int ii, int jj) -> float { return h[(int64_t)ii * ny + jj]; };
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
    float fij, fl, fd, hz, pv = 0.f;
    const bool mid_rows = i >= 3 && i <= nx - 3;  // (i, j), (i-1, j) on no column face
    if (mid_rows && j >= 3 && j <= ny - 3) {      // no face term reaches (i, j), (i, j-1), (i-1, j)
        fij = h[c];
        fl = h[c - 1];
        fd = h[c - ny];
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
    a.lex[off + c] = a.lex2[off + c] + (a.dh / a.dy) * (fl - fij);
    a.ley[off + c] = a.ley2[off + c] + (a.dh / a.dx) * (fij - fd);
    float out = hz + pv;
    if (next_step >= 0) {
        const int p = a.rx_first[c];
        if (p >= 0) out = add_residuals(a, blockIdx.z, c, p, next_step, out);
    }
    a.lhz[off + c] = out;
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
    const int n_steps = a.n_steps > 1 ? a.n_steps : 1;
    if (n_steps > 1 && (a.hist_stride <= 0 || a.step - (n_steps - 1) < 0)) {
        wf::set_error("wf_adjoint_step: %d steps need a positive hist_stride and step >= n_steps-1", n_steps);
        return WF_ERR_ARG;
    }
    const dim3 g((a.ny + kBX - 1) / kBX, (a.nx + kBY - 1) / kBY, a.n_runs);
    wf_adj_args b = a;
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
