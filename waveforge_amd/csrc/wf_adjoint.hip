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
//   k_adj_record  lHz[rx] += dL/dd_n[rx]
//   k_adj_e       transposed Ampere (S9, S8) + coefficient gradients;
//                 lHz gathers the four Ampere neighbours (out of place)
//   k_adj_mur_c   transposed Mur top then bottom (S7, S6), one thread per row i
//   k_adj_mur_r   transposed Mur right then left (S5, S4), one thread per column j
//   k_adj_h       transposed Faraday (S3) gathered per cell, plus the
//                 snapshot's transpose lHz += lP (S1); lP is cleared
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

__global__ void k_adj_record(wf_adj_args a) {
    const int r = blockIdx.x;
    for (int k = threadIdx.x; k < a.n_rx; k += blockDim.x) {
        const int64_t off = (int64_t)r * a.run_stride;
        const float g = a.dl_dd[(int64_t)r * a.dd_run_stride + (int64_t)k * a.dd_rx_stride + a.step];
        atomicAdd(&a.lhz[off + a.rx_cell[k]], g);  // receivers may share a cell
    }
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

__global__ void k_adj_mur_c(wf_adj_args a) {  // S7 then S6 transposed, thread per row i
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nx) return;
    const int ny = a.ny;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride + (int64_t)i * ny;
    float *h = a.lhz2 + off;
    float *p = a.lp + off;
    const float cy = a.cmy;
    float t = h[ny - 1];  // top: Hz[i,ny-1] = P[i,ny-2] + cy (Hz[i,ny-2] - P[i,ny-1])
    p[ny - 2] += t;
    h[ny - 2] += cy * t;
    p[ny - 1] -= cy * t;
    h[ny - 1] = 0.f;
    t = h[0];  // bottom: Hz[i,0] = P[i,1] + cy (Hz[i,1] - P[i,0])
    p[1] += t;
    h[1] += cy * t;
    p[0] -= cy * t;
    h[0] = 0.f;
}

__global__ void k_adj_mur_r(wf_adj_args a) {  // S5 then S4 transposed, thread per column j
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.ny) return;
    const int nx = a.nx, ny = a.ny;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride + j;
    float *h = a.lhz2 + off;
    float *p = a.lp + off;
    const float cx = a.cmx;
    const int64_t r1 = (int64_t)(nx - 1) * ny, r2 = (int64_t)(nx - 2) * ny;
    float t = h[r1];  // right: Hz[nx-1,j] = P[nx-2,j] + cx (Hz[nx-2,j] - P[nx-1,j])
    p[r2] += t;
    h[r2] += cx * t;
    p[r1] -= cx * t;
    h[r1] = 0.f;
    t = h[0];  // left: Hz[0,j] = P[1,j] + cx (Hz[1,j] - P[0,j])
    p[ny] += t;
    h[ny] += cx * t;
    p[0] -= cx * t;
    h[0] = 0.f;
}

__global__ void __launch_bounds__(kBX * kBY) k_adj_h(wf_adj_args a) {
    const int j = blockIdx.x * kBX + threadIdx.x;
    const int i = blockIdx.y * kBY + threadIdx.y;
    const int nx = a.nx, ny = a.ny;
    if (i >= nx || j >= ny) return;
    const int64_t off = (int64_t)blockIdx.z * a.run_stride;
    const int64_t c = (int64_t)i * ny + j;
    const float *h = a.lhz2 + off;
    // Faraday adjoint: Hz[i,j] += Dh ((Ex[i,j+1]-Ex[i,j])/dy - (Ey[i+1,j]-Ey[i,j])/dx), i<=nx-2, j<=ny-2
    auto F = [&](int ii, int jj) -> float {
        return (ii >= 0 && jj >= 0 && ii <= nx - 2 && jj <= ny - 2) ? h[(int64_t)ii * ny + jj] : 0.f;
    };
    const float fij = F(i, j);
    const float lex = a.lex2[off + c] + (a.dh / a.dy) * (F(i, j - 1) - fij);
    const float ley = a.ley2[off + c] + (a.dh / a.dx) * (fij - F(i - 1, j));
    a.lex[off + c] = lex;
    a.ley[off + c] = ley;
    const float pv = a.lp[off + c];
    a.lhz[off + c] = h[c] + pv;
    if (pv != 0.f) a.lp[off + c] = 0.f;
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
        (a.n_rx > 0 && (!a.rx_cell || !a.dl_dd))) {
        wf::set_error("wf_adjoint_step: bad sizes");
        return WF_ERR_ARG;
    }
    if (a.n_rx > 0) hipLaunchKernelGGL(k_adj_record, dim3(a.n_runs), dim3(64), 0, stream, a);
    dim3 g((a.ny + kBX - 1) / kBX, (a.nx + kBY - 1) / kBY, a.n_runs);
    hipLaunchKernelGGL(k_adj_e, g, dim3(kBX, kBY), 0, stream, a);
    hipLaunchKernelGGL(k_adj_mur_c, dim3((a.nx + 63) / 64, 1, a.n_runs), dim3(64), 0, stream, a);
    hipLaunchKernelGGL(k_adj_mur_r, dim3((a.ny + 63) / 64, 1, a.n_runs), dim3(64), 0, stream, a);
    hipLaunchKernelGGL(k_adj_h, g, dim3(kBX, kBY), 0, stream, a);
    return wf::check_launch("wf_adjoint_step");
}
