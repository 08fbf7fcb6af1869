// Reference-exact 3D Yee step (FDTD3D.step, src/core/fdtd3d.py:262-373, with
// MurABC3D, src/core/boundaries.py:461-575), in place, one launch per stage:
//
//   k3_snap   copy the H values Mur reads from its snapshot: for each
//             component, the planes at index 0, 1, n-2, n-1 across each of
//             its two tangential axes (Hy: x and z, Hz: x and y, Hx: y and z);
//   k3_src    soft sources in collection order, one thread per (component,
//             cell) group (src/core/sources.py:1060-1071);
//   k3_h      Faraday for Hx on [:, :-1, :-1], Hy on [:-1, :, :-1], Hz on
//             [:-1, :-1, :] (fdtd3d.py:309-330);
//   k3_mur1   Mur across each component's first tangential axis (Hy and Hz
//             on the x faces, Hx on the y faces);
//   k3_mur2   Mur across the second (Hz on the y faces; Hx and Hy on the z
//             faces when nz > 1). The reference applies the faces in the order
//             x, y, z, so a second-axis face reads first-axis faces already
//             updated (edges), which the launch order reproduces;
//   k3_e      Ampere for Ex on [:, 1:, 1:], Ey on [1:, :, 1:], Ez on [1:, 1:, :]
//             (fdtd3d.py:336-359);
//   k3_rec    detectors (any component) after the step.
//
// Layout: (nx, ny, nz) row-major, k fastest: cell = (i*ny + j)*nz + k. All
// arithmetic is fp32 in the reference's operation order (-ffp-contract=off,
// IEEE division by fp32(d)).
#include "wf_common.hpp"

#include <algorithm>

namespace {

constexpr int kBX = 64;
constexpr int kBY = 4;

__device__ __forceinline__ float *comp_ptr(const wf_run3d &r, int c) {
    switch (c) {
        case WF_COMP_EX: return r.ex;
        case WF_COMP_EY: return r.ey;
        case WF_COMP_EZ: return r.ez;
        case WF_COMP_HX: return r.hx;
        case WF_COMP_HY: return r.hy;
        default: return r.hz;
    }
}

__device__ __forceinline__ int plane_index(int s, int n) { return s < 2 ? s : n - 4 + s; }

// Snapshot plane sets (see wf_fdtd3d_strip_floats): A Hy@x, B Hz@x (ny*nz each
// plane), C Hx@y, D Hz@y (nx*nz), E Hx@z, F Hy@z (nx*ny, only when nz > 1).
struct Strips {
    int64_t a, b, c, d, e, f, total;
    __host__ __device__ Strips(int nx, int ny, int nz) {
        const int64_t yz = (int64_t)ny * nz, xz = (int64_t)nx * nz, xy = (int64_t)nx * ny;
        a = 0;
        b = a + 4 * yz;
        c = b + 4 * yz;
        d = c + 4 * xz;
        e = d + 4 * xz;
        f = e + (nz > 1 ? 4 * xy : 0);
        total = f + (nz > 1 ? 4 * xy : 0);
    }
};

__global__ void __launch_bounds__(256) k3_snap(wf_run3d r) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const Strips S(nx, ny, nz);
    const int64_t yz = (int64_t)ny * nz, xz = (int64_t)nx * nz, xy = (int64_t)nx * ny;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < S.total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t src;
        const float *f;
        if (t < S.c) {  // A, B: planes across x
            const int64_t u = t < S.b ? t : t - S.b;
            const int s = (int)(u / yz);
            const int64_t jk = u - s * yz;
            src = (int64_t)plane_index(s, nx) * yz + jk;
            f = t < S.b ? r.hy : r.hz;
        } else if (t < S.e) {  // C, D: planes across y
            const int64_t u = t < S.d ? t - S.c : t - S.d;
            const int s = (int)(u / xz);
            const int64_t ik = u - s * xz;
            const int64_t i = ik / nz, k = ik - i * nz;
            src = (i * ny + plane_index(s, ny)) * nz + k;
            f = t < S.d ? r.hx : r.hz;
        } else {  // E, F: planes across z
            const int64_t u = t < S.f ? t - S.e : t - S.f;
            const int s = (int)(u / xy);
            const int64_t ij = u - s * xy;
            src = ij * nz + plane_index(s, nz);
            f = t < S.f ? r.hx : r.hy;
        }
        r.pstrip[t] = f[src];
    }
}

__global__ void __launch_bounds__(256) k3_src(wf_run3d r, int n) {
    for (int g = threadIdx.x; g < r.n_grp; g += blockDim.x) {
        const int a = r.grp_start[g], b = r.grp_start[g + 1];
        float *f = comp_ptr(r, r.inj[a].comp);
        const int64_t cell = r.inj[a].cell;
        float v = f[cell];
        for (int c = a; c < b; ++c) {
            const float w = r.inj[c].wave[n] * r.inj[c].amp;  // values = w_now * amp
            v = v + w;
        }
        f[cell] = v;
    }
}

__global__ void __launch_bounds__(kBX * kBY) k3_h(wf_run3d r) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= nz || j >= ny) return;
    const int64_t sj = nz, si = (int64_t)ny * nz;
    const int64_t c = (int64_t)i * si + (int64_t)j * sj + k;
    if (j <= ny - 2 && k <= nz - 2) {
        const float a = (r.ey[c + 1] - r.ey[c]) / r.dz;
        const float b = (r.ez[c + sj] - r.ez[c]) / r.dy;
        r.hx[c] = r.hx[c] + r.dh * (a - b);
    }
    if (i <= nx - 2 && k <= nz - 2) {
        const float a = (r.ez[c + si] - r.ez[c]) / r.dx;
        const float b = (r.ex[c + 1] - r.ex[c]) / r.dz;
        r.hy[c] = r.hy[c] + r.dh * (a - b);
    }
    if (i <= nx - 2 && j <= ny - 2) {
        const float a = (r.ex[c + sj] - r.ex[c]) / r.dy;
        const float b = (r.ey[c + si] - r.ey[c]) / r.dx;
        r.hz[c] = r.hz[c] + r.dh * (a - b);
    }
}

// f[face] = P[in] + C * (f[in] - P[face]) for the min (s=0) or max (s=1) face.
__device__ __forceinline__ void mur_face(float *f, int64_t face, int64_t in, float p_in, float p_face, float C) {
    f[face] = p_in + C * (f[in] - p_face);
}

__global__ void __launch_bounds__(256) k3_mur1(wf_run3d r) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const Strips S(nx, ny, nz);
    const int64_t yz = (int64_t)ny * nz, xz = (int64_t)nx * nz;
    const int64_t n_x = 2 * yz, n_y = 2 * xz;  // cells per component: two faces
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * n_x) {  // Hy then Hz on x faces
        const bool is_hz = t >= n_x;
        const int64_t u = is_hz ? t - n_x : t;
        const int side = (int)(u / yz);
        const int64_t jk = u - side * yz;
        const float *P = r.pstrip + (is_hz ? S.b : S.a);
        float *f = is_hz ? r.hz : r.hy;
        const int fi = side ? nx - 1 : 0, ii = side ? nx - 2 : 1;
        const float p_in = P[(side ? 2 : 1) * yz + jk], p_face = P[(side ? 3 : 0) * yz + jk];
        mur_face(f, fi * yz + jk, ii * yz + jk, p_in, p_face, r.cmx);
    } else if (t < 2 * n_x + n_y) {  // Hx on y faces
        const int64_t u = t - 2 * n_x;
        const int side = (int)(u / xz);
        const int64_t ik = u - side * xz;
        const int64_t i = ik / nz, k = ik - i * nz;
        const float *P = r.pstrip + S.c;
        const int fj = side ? ny - 1 : 0, ij = side ? ny - 2 : 1;
        const float p_in = P[(side ? 2 : 1) * xz + ik], p_face = P[(side ? 3 : 0) * xz + ik];
        mur_face(r.hx, (i * ny + fj) * nz + k, (i * ny + ij) * nz + k, p_in, p_face, r.cmy);
    }
}

__global__ void __launch_bounds__(256) k3_mur2(wf_run3d r) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const Strips S(nx, ny, nz);
    const int64_t xz = (int64_t)nx * nz, xy = (int64_t)nx * ny;
    const int64_t n_y = 2 * xz, n_z = nz > 1 ? 2 * xy : 0;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_y) {  // Hz on y faces
        const int side = (int)(t / xz);
        const int64_t ik = t - side * xz;
        const int64_t i = ik / nz, k = ik - i * nz;
        const float *P = r.pstrip + S.d;
        const int fj = side ? ny - 1 : 0, ij = side ? ny - 2 : 1;
        const float p_in = P[(side ? 2 : 1) * xz + ik], p_face = P[(side ? 3 : 0) * xz + ik];
        mur_face(r.hz, (i * ny + fj) * nz + k, (i * ny + ij) * nz + k, p_in, p_face, r.cmy);
    } else if (t < n_y + 2 * n_z) {  // Hx then Hy on z faces
        const bool is_hy = t >= n_y + n_z;
        const int64_t u = t - n_y - (is_hy ? n_z : 0);
        const int side = (int)(u / xy);
        const int64_t ij = u - side * xy;
        const float *P = r.pstrip + (is_hy ? S.f : S.e);
        float *f = is_hy ? r.hy : r.hx;
        const int fk = side ? nz - 1 : 0, ik = side ? nz - 2 : 1;
        const float p_in = P[(side ? 2 : 1) * xy + ij], p_face = P[(side ? 3 : 0) * xy + ij];
        mur_face(f, ij * nz + fk, ij * nz + ik, p_in, p_face, r.cmz);
    }
}

template <bool kMat>
__global__ void __launch_bounds__(kBX * kBY) k3_e(wf_run3d r) {
    const int ny = r.ny, nz = r.nz;
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= nz || j >= ny) return;
    const int64_t sj = nz, si = (int64_t)ny * nz;
    const int64_t c = (int64_t)i * si + (int64_t)j * sj + k;
    if (j >= 1 && k >= 1) {
        const float a = (r.hz[c] - r.hz[c - sj]) / r.dy;
        const float b = (r.hy[c] - r.hy[c - 1]) / r.dz;
        const float curl = a - b;
        if (kMat) {
            const float t1 = r.ca[c] * r.ex[c];
            const float t2 = r.cb[c] * curl;
            r.ex[c] = t1 + t2;
        } else {
            r.ex[c] = r.ex[c] + r.de * curl;
        }
    }
    if (i >= 1 && k >= 1) {
        const float a = (r.hx[c] - r.hx[c - 1]) / r.dz;
        const float b = (r.hz[c] - r.hz[c - si]) / r.dx;
        const float curl = a - b;
        if (kMat) {
            const float t1 = r.ca[c] * r.ey[c];
            const float t2 = r.cb[c] * curl;
            r.ey[c] = t1 + t2;
        } else {
            r.ey[c] = r.ey[c] + r.de * curl;
        }
    }
    if (i >= 1 && j >= 1) {
        const float a = (r.hy[c] - r.hy[c - si]) / r.dx;
        const float b = (r.hx[c] - r.hx[c - sj]) / r.dy;
        const float curl = a - b;
        if (kMat) {
            const float t1 = r.ca[c] * r.ez[c];
            const float t2 = r.cb[c] * curl;
            r.ez[c] = t1 + t2;
        } else {
            r.ez[c] = r.ez[c] + r.de * curl;
        }
    }
}

__global__ void __launch_bounds__(256) k3_rec(wf_run3d r, int n) {
    const int t = n - r.trace_n0;
    if (t < 0 || t >= r.trace_len) return;
    for (int d = threadIdx.x; d < r.n_det; d += blockDim.x)
        r.traces[d * r.trace_stride + t] = comp_ptr(r, r.det_comp[d])[r.det_cell[d]];
}

int validate3(const wf_run3d &r) {
    if (!r.ex || !r.ey || !r.ez || !r.hx || !r.hy || !r.hz || !r.pstrip) {
        wf::set_error("run3d: null field or strip pointer");
        return WF_ERR_ARG;
    }
    if (r.nx < 4 || r.ny < 4 || (r.nz != 1 && r.nz < 4)) {
        wf::set_error("run3d: grid must be at least 4x4 with nz = 1 or nz >= 4, got %dx%dx%d", r.nx, r.ny, r.nz);
        return WF_ERR_ARG;
    }
    if ((r.ca == nullptr) != (r.cb == nullptr)) {
        wf::set_error("run3d: ca and cb must both be set or both NULL");
        return WF_ERR_ARG;
    }
    if (r.n_grp < 0 || r.n_det < 0 || (r.n_grp > 0 && (!r.inj || !r.grp_start)) ||
        (r.n_det > 0 && (!r.det_cell || !r.det_comp || !r.traces))) {
        wf::set_error("run3d: source or detector tables missing");
        return WF_ERR_ARG;
    }
    return WF_OK;
}

}  // namespace

extern "C" int64_t wf_fdtd3d_strip_floats(int32_t nx, int32_t ny, int32_t nz) { return Strips(nx, ny, nz).total; }

extern "C" int wf_fdtd3d_step(const wf_run3d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream) {
    if (!runs || n_runs < 1 || n_steps < 0) {
        wf::set_error("wf_fdtd3d_step: bad arguments");
        return WF_ERR_ARG;
    }
    for (int b = 0; b < n_runs; ++b) {
        const int rc = validate3(runs[b]);
        if (rc) return rc;
    }
    for (int s = 0; s < n_steps; ++s) {
        for (int b = 0; b < n_runs; ++b) {
            const wf_run3d &r = runs[b];
            const int n = r.step + s;
            const Strips S(r.nx, r.ny, r.nz);
            const unsigned snap_blocks = (unsigned)std::min<int64_t>((S.total + 255) / 256, 4096);
            hipLaunchKernelGGL(k3_snap, dim3(snap_blocks), dim3(256), 0, stream, r);
            if (r.n_grp > 0) hipLaunchKernelGGL(k3_src, dim3(1), dim3(256), 0, stream, r, n);
            const dim3 g3((r.nz + kBX - 1) / kBX, (r.ny + kBY - 1) / kBY, r.nx);
            hipLaunchKernelGGL(k3_h, g3, dim3(kBX, kBY), 0, stream, r);
            const int64_t m1 = 4 * (int64_t)r.ny * r.nz + 2 * (int64_t)r.nx * r.nz;
            hipLaunchKernelGGL(k3_mur1, dim3((unsigned)((m1 + 255) / 256)), dim3(256), 0, stream, r);
            const int64_t m2 = 2 * (int64_t)r.nx * r.nz + (r.nz > 1 ? 4 * (int64_t)r.nx * r.ny : 0);
            hipLaunchKernelGGL(k3_mur2, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, stream, r);
            if (r.ca)
                hipLaunchKernelGGL(k3_e<true>, g3, dim3(kBX, kBY), 0, stream, r);
            else
                hipLaunchKernelGGL(k3_e<false>, g3, dim3(kBX, kBY), 0, stream, r);
            if (r.n_det > 0) hipLaunchKernelGGL(k3_rec, dim3(1), dim3(256), 0, stream, r, n);
        }
        const int rc = wf::check_launch("wf_fdtd3d_step");
        if (rc) return rc;
    }
    return WF_OK;
}
