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

// ---------------------------------------------------------------------------
// Fused path: one streaming pass per step over ping-pong buffers.
//
// A workgroup of 16 x 64 threads owns a (j, k) tile: rows j0-1 .. j0+14 and
// lanes k0-1 .. k0+62, i.e. a core of 14 x 62 cells with a one-cell halo, and
// walks i over a chunk of planes (starting one plane early so the
// previous plane's new H is in registers). For each plane it stages the old E
// of the plane in LDS, computes Faraday (new H, "F") on every cell but the
// high-side halo from the LDS neighbours j+1, k+1 and the next plane's E in
// registers, stages F in LDS, and writes the core's F and, on the cells whose
// Ampere stencil touches no Mur face cell ([2, n-2] on every axis), the new
// E from the LDS neighbours j-1, k-1 and the previous plane's F. The Mur faces
// (k3_mur1/k3_mur2 on the written state) and the E cells outside that box
// (k3_eshell) follow. Reads come only from the old state, so tiles and chunks
// never race.
// ---------------------------------------------------------------------------
constexpr int kFLanes = 64;
constexpr int kFRows = 16;
constexpr int kFCoreK = kFLanes - 2;
constexpr int kFCoreJ = kFRows - 2;
constexpr int kResidentBlocks = 512;  // 256 CUs x 2 workgroups of 1024 threads

struct Bufs {
    const float *ex, *ey, *ez, *hx, *hy, *hz;  // state read (old)
    float *nex, *ney, *nez, *nhx, *nhy, *nhz;  // state written (new)
};

template <bool kMat>
__global__ void __launch_bounds__(kFLanes * kFRows) __attribute__((amdgpu_waves_per_eu(8, 8)))
k3_fused(wf_run3d r, Bufs b, int chunk, int ntk, int ntj, int n_tiles) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int l = threadIdx.x, row = threadIdx.y;
    // XCD-aware tile order: workgroup b runs on XCD b % 8; each XCD walks a
    // contiguous band of tiles (k fastest, then j, then i-chunk), so the halo
    // rows and misaligned row ends that neighbouring tiles share meet in one L2
    const int per = (n_tiles + 7) / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= n_tiles) return;
    const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
    const int k = tk * kFCoreK - 1 + l;
    const int j = tj * kFCoreJ - 1 + row;
    const int i0 = ti * chunk;
    const int i1 = min(nx, i0 + chunk);
    const bool in_g = j >= 0 && j < ny && k >= 0 && k < nz;
    const bool f_ok = in_g && row <= kFRows - 2 && l <= kFLanes - 2;
    const bool core = in_g && row >= 1 && row <= kFRows - 2 && l >= 1 && l <= kFLanes - 2;
    // With nz >= 4 the z faces of Hx, Hy on rows j in [1, ny-2] of planes
    // i in [1, nx-2] are finished here (their Mur reads no x- or y-face value),
    // so Ampere runs on every k of the interior (i, j) columns; with nz = 1
    // (no z faces) it runs on the [2, n-2] box and the shell kernel does the rest.
    const bool zext = nz >= 4;
    const bool box_jk = core && j >= 2 && j <= ny - 2 && (zext || (k >= 2 && k <= nz - 2));
    const bool zface = zext && f_ok && (k == 0 || k == nz - 1) && j >= 1 && j <= ny - 2;
    const Strips S(nx, ny, nz);
    const int64_t xy = (int64_t)nx * ny;
    const int64_t si = (int64_t)ny * nz, sj = nz;
    const int64_t cjk = in_g ? (int64_t)j * sj + k : 0;
    __shared__ float sE[3][kFRows][kFLanes];
    __shared__ float sF[3][kFRows][kFLanes];
    int i = i0 > 0 ? i0 - 1 : 0;
    // registers: E of planes i and i+1 and H of plane i; the E and H loads for
    // the following plane are issued before the barriers
    float ex = 0.f, ey = 0.f, ez = 0.f, exn = 0.f, eyn = 0.f, ezn = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    if (in_g) {
        const int64_t c = (int64_t)i * si + cjk;
        ex = b.ex[c];
        ey = b.ey[c];
        ez = b.ez[c];
        if (i + 1 < nx) {
            exn = b.ex[c + si];
            eyn = b.ey[c + si];
            ezn = b.ez[c + si];
        }
        hx = b.hx[c];
        hy = b.hy[c];
        hz = b.hz[c];
    }
    float fhy_prev = 0.f, fhz_prev = 0.f;
    for (; i < i1; ++i) {
        const int64_t c = (int64_t)i * si + cjk;
        float exnn = 0.f, eynn = 0.f, eznn = 0.f, hxn = 0.f, hyn = 0.f, hzn = 0.f, ca = 0.f, cb = 0.f;
        if (kMat && box_jk && i >= i0) {  // used after two barriers: loaded for this plane
            ca = r.ca[c];
            cb = r.cb[c];
        }
        if (in_g && i + 1 < i1) {
            if (i + 2 < nx) {
                exnn = b.ex[c + 2 * si];
                eynn = b.ey[c + 2 * si];
                eznn = b.ez[c + 2 * si];
            }
            hxn = b.hx[c + si];
            hyn = b.hy[c + si];
            hzn = b.hz[c + si];
        }
        sE[0][row][l] = ex;
        sE[1][row][l] = ey;
        sE[2][row][l] = ez;
        __syncthreads();
        float fhx = hx, fhy = hy, fhz = hz;
        if (f_ok) {
            if (j <= ny - 2 && k <= nz - 2) {
                const float a = (sE[1][row][l + 1] - ey) / r.dz;
                const float bb = (sE[2][row + 1][l] - ez) / r.dy;
                fhx = hx + r.dh * (a - bb);
            }
            if (i <= nx - 2 && k <= nz - 2) {
                const float a = (ezn - ez) / r.dx;
                const float bb = (sE[0][row][l + 1] - ex) / r.dz;
                fhy = hy + r.dh * (a - bb);
            }
            if (i <= nx - 2 && j <= ny - 2) {
                const float a = (sE[0][row + 1][l] - ex) / r.dy;
                const float bb = (eyn - ey) / r.dx;
                fhz = hz + r.dh * (a - bb);
            }
        }
        sF[0][row][l] = fhx;
        sF[1][row][l] = fhy;
        sF[2][row][l] = fhz;
        __syncthreads();
        if (zface && i >= 1 && i <= nx - 2) {
            // Mur on the z faces (boundaries.py:557-575): f[face] = P[in] + cz (f[in] - P[face]),
            // the inner value from the neighbouring lane, P from the snapshot planes
            const int64_t ij = (int64_t)i * ny + j;
            const float *PE = r.pstrip + S.e, *PF = r.pstrip + S.f;
            const int lin = k == 0 ? l + 1 : l - 1;
            const int s_in = k == 0 ? 1 : 2, s_face = k == 0 ? 0 : 3;
            fhx = PE[s_in * xy + ij] + r.cmz * (sF[0][row][lin] - PE[s_face * xy + ij]);
            fhy = PF[s_in * xy + ij] + r.cmz * (sF[1][row][lin] - PF[s_face * xy + ij]);
            sF[0][row][l] = fhx;
            sF[1][row][l] = fhy;
        }
        __syncthreads();
        if (core && i >= i0) {
            b.nhx[c] = fhx;
            b.nhy[c] = fhy;
            b.nhz[c] = fhz;
            if (box_jk && i >= 2 && i <= nx - 2) {
                float vx = ex, vy = ey;
                if (k >= 1) {
                    const float cx = ((fhz - sF[2][row - 1][l]) / r.dy) - ((fhy - sF[1][row][l - 1]) / r.dz);
                    const float cy = ((fhx - sF[0][row][l - 1]) / r.dz) - ((fhz - fhz_prev) / r.dx);
                    vx = kMat ? ca * ex + cb * cx : ex + r.de * cx;
                    vy = kMat ? ca * ey + cb * cy : ey + r.de * cy;
                }
                const float cz = ((fhy - fhy_prev) / r.dx) - ((fhx - sF[0][row - 1][l]) / r.dy);
                b.nex[c] = vx;
                b.ney[c] = vy;
                b.nez[c] = kMat ? ca * ez + cb * cz : ez + r.de * cz;
            }
        }
        fhy_prev = fhy;
        fhz_prev = fhz;
        ex = exn;
        ey = eyn;
        ez = ezn;
        exn = exnn;
        eyn = eynn;
        ezn = eznn;
        hx = hxn;
        hy = hyn;
        hz = hzn;
    }
}

// Shell of an axis of length n: the indices outside [2, n-2] ({0, 1, n-1} for
// n >= 4, {0} for n = 1), and the count of indices inside.
__device__ __forceinline__ int shell_count(int n) { return n == 1 ? 1 : 3; }
__device__ __forceinline__ int shell_index(int t, int n) { return t < 2 ? t : n - 1; }

// E on every cell the fused kernel leaves: the planes i in {0, 1, nx-1}, the
// rows j in {0, 1, ny-1} of the other planes, and (nz = 1 only) the rest;
// updated where the reference updates it (from the final new H), copied where
// it never does.
template <bool kMat>
__global__ void __launch_bounds__(256) k3_eshell(wf_run3d r, Bufs b) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int sx = shell_count(nx), sy = shell_count(ny), sz = shell_count(nz);
    const int mx = nx - sx, my = ny - sy;
    const int64_t n1 = (int64_t)sx * ny * nz, n2 = (int64_t)mx * sy * nz;
    const int64_t n3 = nz == 1 ? (int64_t)mx * my * sz : 0;  // with nz >= 4 the fused kernel does the k-slab
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < n1) {
        const int64_t yzc = (int64_t)ny * nz;
        i = shell_index((int)(t / yzc), nx);
        const int64_t u = t % yzc;
        j = (int)(u / nz);
        k = (int)(u % nz);
    } else if ((t -= n1) < n2) {
        const int64_t per_i = (int64_t)sy * nz;
        i = 2 + (int)(t / per_i);
        const int64_t u = t % per_i;
        j = shell_index((int)(u / nz), ny);
        k = (int)(u % nz);
    } else if ((t -= n2) < n3) {
        const int64_t per_i = (int64_t)my * sz;
        i = 2 + (int)(t / per_i);
        const int64_t u = t % per_i;
        j = 2 + (int)(u / sz);
        k = shell_index((int)(u % sz), nz);
    } else {
        return;
    }
    const int64_t sj = nz, si = (int64_t)ny * nz;
    const int64_t c = (int64_t)i * si + (int64_t)j * sj + k;
    const float *hx = b.nhx, *hy = b.nhy, *hz = b.nhz;
    float vx = b.ex[c], vy = b.ey[c], vz = b.ez[c];
    if (j >= 1 && k >= 1) {
        const float curl = ((hz[c] - hz[c - sj]) / r.dy) - ((hy[c] - hy[c - 1]) / r.dz);
        vx = kMat ? r.ca[c] * vx + r.cb[c] * curl : vx + r.de * curl;
    }
    if (i >= 1 && k >= 1) {
        const float curl = ((hx[c] - hx[c - 1]) / r.dz) - ((hz[c] - hz[c - si]) / r.dx);
        vy = kMat ? r.ca[c] * vy + r.cb[c] * curl : vy + r.de * curl;
    }
    if (i >= 1 && j >= 1) {
        const float curl = ((hy[c] - hy[c - si]) / r.dx) - ((hx[c] - hx[c - sj]) / r.dy);
        vz = kMat ? r.ca[c] * vz + r.cb[c] * curl : vz + r.de * curl;
    }
    b.nex[c] = vx;
    b.ney[c] = vy;
    b.nez[c] = vz;
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

wf_run3d view_of(const wf_run3d &r, bool second) {
    wf_run3d v = r;
    if (second) {
        v.ex = r.ex2; v.ey = r.ey2; v.ez = r.ez2;
        v.hx = r.hx2; v.hy = r.hy2; v.hz = r.hz2;
    }
    return v;
}

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

extern "C" int wf_fdtd3d_fused_step(const wf_run3d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream) {
    if (!runs || n_runs < 1 || n_steps < 0 || (n_steps & 1)) {
        wf::set_error("wf_fdtd3d_fused_step: bad arguments (n_steps must be even)");
        return WF_ERR_ARG;
    }
    for (int b = 0; b < n_runs; ++b) {
        const int rc = validate3(runs[b]);
        if (rc) return rc;
        const wf_run3d &r = runs[b];
        if (!r.ex2 || !r.ey2 || !r.ez2 || !r.hx2 || !r.hy2 || !r.hz2) {
            wf::set_error("wf_fdtd3d_fused_step: second state buffers missing");
            return WF_ERR_ARG;
        }
    }
    for (int s = 0; s < n_steps; ++s) {
        const bool odd = s & 1;
        for (int bi = 0; bi < n_runs; ++bi) {
            const wf_run3d rd = view_of(runs[bi], odd);   // read state
            const wf_run3d wr = view_of(runs[bi], !odd);  // written state
            const int n = runs[bi].step + s;
            const Strips S(rd.nx, rd.ny, rd.nz);
            const unsigned snap_blocks = (unsigned)std::min<int64_t>((S.total + 255) / 256, 4096);
            hipLaunchKernelGGL(k3_snap, dim3(snap_blocks), dim3(256), 0, stream, rd);
            if (rd.n_grp > 0) hipLaunchKernelGGL(k3_src, dim3(1), dim3(256), 0, stream, rd, n);
            Bufs bf{rd.ex, rd.ey, rd.ez, rd.hx, rd.hy, rd.hz, wr.ex, wr.ey, wr.ez, wr.hx, wr.hy, wr.hz};
            // i-chunks sized so the grid is about one round of resident workgroups
            const int tiles = ((rd.nz + kFCoreK - 1) / kFCoreK) * ((rd.ny + kFCoreJ - 1) / kFCoreJ);
            const int n_chunks = std::max(1, std::min((rd.nx + 7) / 8, kResidentBlocks / std::max(1, tiles)));
            const int chunk = (rd.nx + n_chunks - 1) / n_chunks;
            const int ntk = (rd.nz + kFCoreK - 1) / kFCoreK, ntj = (rd.ny + kFCoreJ - 1) / kFCoreJ;
            const int n_tiles = ntk * ntj * ((rd.nx + chunk - 1) / chunk);
            const dim3 gf(8u * (unsigned)((n_tiles + 7) / 8));
            if (rd.ca)
                hipLaunchKernelGGL(k3_fused<true>, gf, dim3(kFLanes, kFRows), 0, stream, rd, bf, chunk, ntk, ntj,
                                   n_tiles);
            else
                hipLaunchKernelGGL(k3_fused<false>, gf, dim3(kFLanes, kFRows), 0, stream, rd, bf, chunk, ntk, ntj,
                                   n_tiles);
            const int64_t m1 = 4 * (int64_t)wr.ny * wr.nz + 2 * (int64_t)wr.nx * wr.nz;
            hipLaunchKernelGGL(k3_mur1, dim3((unsigned)((m1 + 255) / 256)), dim3(256), 0, stream, wr);
            const int64_t m2 = 2 * (int64_t)wr.nx * wr.nz + (wr.nz > 1 ? 4 * (int64_t)wr.nx * wr.ny : 0);
            hipLaunchKernelGGL(k3_mur2, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, stream, wr);
            const int sx = wr.nx == 1 ? 1 : 3, sy = wr.ny == 1 ? 1 : 3, sz = wr.nz == 1 ? 1 : 3;
            const int64_t nsh = (int64_t)sx * wr.ny * wr.nz + (int64_t)(wr.nx - sx) * sy * wr.nz +
                                (wr.nz == 1 ? (int64_t)(wr.nx - sx) * (wr.ny - sy) * sz : 0);
            if (rd.ca)
                hipLaunchKernelGGL(k3_eshell<true>, dim3((unsigned)((nsh + 255) / 256)), dim3(256), 0, stream, rd, bf);
            else
                hipLaunchKernelGGL(k3_eshell<false>, dim3((unsigned)((nsh + 255) / 256)), dim3(256), 0, stream, rd, bf);
            if (wr.n_det > 0) hipLaunchKernelGGL(k3_rec, dim3(1), dim3(256), 0, stream, wr, n);
        }
        const int rc = wf::check_launch("wf_fdtd3d_fused_step");
        if (rc) return rc;
    }
    return WF_OK;
}
