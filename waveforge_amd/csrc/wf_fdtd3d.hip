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
            if ((unsigned)n >= (unsigned)r.inj[c].wave_len) {  // past the table: report, inject nothing
                if (r.status) atomicOr(r.status, WF_STATUS_WAVE_END);
                continue;
            }
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

// The H of the read state: where the snapshot is taken from when the run
// carries WF_RUN3D_SNAP_FROM_STATE (the fused path reads one state and writes
// another, so the read state still holds the step's starting H).
struct Prev {
    const float *hx, *hy, *hz;
};

// Snapshot value of component f at plane index s (0, 1, n-2, n-1 for s = 0..3)
// across an axis: from the strip planes at P[s*plane + idx], or from the read
// state at `cell`.
template <bool kState>
__device__ __forceinline__ float snap(const float *P, int s, int64_t plane, int64_t idx, const float *f, int64_t cell) {
    return kState ? f[cell] : P[s * plane + idx];
}

template <bool kState>
__global__ void __launch_bounds__(256) k3_mur1(wf_run3d r, Prev pv) {
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
        const float *F = is_hz ? pv.hz : pv.hy;
        float *f = is_hz ? r.hz : r.hy;
        const int fi = side ? nx - 1 : 0, ii = side ? nx - 2 : 1;
        const int64_t face = fi * yz + jk, in = ii * yz + jk;
        mur_face(f, face, in, snap<kState>(P, side ? 2 : 1, yz, jk, F, in), snap<kState>(P, side ? 3 : 0, yz, jk, F, face),
                 r.cmx);
    } else if (t < 2 * n_x + n_y) {  // Hx on y faces
        const int64_t u = t - 2 * n_x;
        const int side = (int)(u / xz);
        const int64_t ik = u - side * xz;
        const int64_t i = ik / nz, k = ik - i * nz;
        const float *P = r.pstrip + S.c;
        const int fj = side ? ny - 1 : 0, ij = side ? ny - 2 : 1;
        const int64_t face = (i * ny + fj) * nz + k, in = (i * ny + ij) * nz + k;
        mur_face(r.hx, face, in, snap<kState>(P, side ? 2 : 1, xz, ik, pv.hx, in),
                 snap<kState>(P, side ? 3 : 0, xz, ik, pv.hx, face), r.cmy);
    }
}

// z-face columns (i, j) per face: all nx*ny, or (kEdges) only those on an x or
// y face, which are the ones the fused pass leaves when it applies the rest.
template <bool kEdges>
__host__ __device__ __forceinline__ int64_t zface_columns(int nx, int ny, int nz) {
    if (nz == 1) return 0;
    return kEdges ? 2 * (int64_t)ny + 2 * (int64_t)(nx - 2) : (int64_t)nx * ny;
}

template <bool kState>
__global__ void __launch_bounds__(256) k3_mur2(wf_run3d r, Prev pv) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const Strips S(nx, ny, nz);
    const int64_t xz = (int64_t)nx * nz, xy = (int64_t)nx * ny;
    const int64_t zc = zface_columns<kState>(nx, ny, nz);
    const int64_t n_y = 2 * xz, n_z = 2 * zc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_y) {  // Hz on y faces
        const int side = (int)(t / xz);
        const int64_t ik = t - side * xz;
        const int64_t i = ik / nz, k = ik - i * nz;
        const float *P = r.pstrip + S.d;
        const int fj = side ? ny - 1 : 0, ij = side ? ny - 2 : 1;
        const int64_t face = (i * ny + fj) * nz + k, in = (i * ny + ij) * nz + k;
        mur_face(r.hz, face, in, snap<kState>(P, side ? 2 : 1, xz, ik, pv.hz, in),
                 snap<kState>(P, side ? 3 : 0, xz, ik, pv.hz, face), r.cmy);
    } else if (t < n_y + 2 * n_z) {  // Hx then Hy on z faces
        const bool is_hy = t >= n_y + n_z;
        const int64_t u = t - n_y - (is_hy ? n_z : 0);
        const int side = (int)(u / zc);
        int64_t ij = u - side * zc;
        if (kState) {  // edge columns: i in {0, nx-1} (every j), then j in {0, ny-1} (i in [1, nx-2])
            const int64_t e = ij;
            if (e < 2 * (int64_t)ny) {
                ij = (e < ny ? 0 : (int64_t)(nx - 1) * ny) + e % ny;
            } else {
                const int64_t f = e - 2 * (int64_t)ny;
                ij = (1 + f % (nx - 2)) * ny + (f < nx - 2 ? 0 : ny - 1);
            }
        }
        const float *P = r.pstrip + (is_hy ? S.f : S.e);
        const float *F = is_hy ? pv.hy : pv.hx;
        float *f = is_hy ? r.hy : r.hx;
        const int fk = side ? nz - 1 : 0, ik = side ? nz - 2 : 1;
        const int64_t face = ij * nz + fk, in = ij * nz + ik;
        mur_face(f, face, in, snap<kState>(P, side ? 2 : 1, xy, ij, F, in), snap<kState>(P, side ? 3 : 0, xy, ij, F, face),
                 r.cmz);
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
// A wavefront owns a (j, k) tile of R rows (j0 .. j0+R-1) by 64 lanes
// (k0 .. k0+63, k0 a multiple of 64, so each row access is whole aligned
// lines) and walks i over a chunk of planes, starting one plane early so the
// previous plane's new H is in registers. Row j0-1 (whose new Hx, Hz Ampere
// reads at j-1) and row j0+R (whose old Ex, Ez Faraday reads at j+1) ride
// along. The k-neighbours are the adjacent lanes (DPP whole-wave shifts); the
// two that cross the tile's k edges -- old Ex, Ey at k0+64 for Faraday in
// lane 63 and the new Hx, Hy at k0-1 for Ampere in lane 0 -- are loaded and
// computed once per row in spare lanes (lane q serves row q) and handed to
// the edge lane with readlane. There is no LDS and no barrier: every
// wavefront is independent; a workgroup holds consecutive tiles only so that
// they run side by side on one CU.
//
// Per plane: Faraday on every tile cell; with WF_RUN3D_SNAP_FROM_STATE the Mur
// z faces of Hx, Hy on rows j in [1, ny-2] of planes i in [1, nx-2] (they read
// no x- or y-face value; the snapshot is the old H in registers); then the new
// H of the core rows and, on the cells whose Ampere stencil touches no Mur face
// cell (i, j in [2, n-2]; k in [2, nz-2], or every k when the z faces were
// applied here), the new E. The x and y faces (k3_mur1/k3_mur2 on the written
// state; z faces again on the edge columns) and the E cells outside that box
// (k3_eshell) follow. Reads come only from the old state, so tiles and chunks
// never race.
// ---------------------------------------------------------------------------
constexpr int kSRows = 4;   // rows per wavefront tile
constexpr int kSWaves = 4;  // wavefronts per SIMD the register budget is sized for
constexpr int kSGroup = 4;  // wavefronts per workgroup (consecutive tiles)

struct Bufs {
    const float *ex, *ey, *ez, *hx, *hy, *hz;  // state read (old)
    float *nex, *ney, *nez, *nhx, *nhy, *nhz;  // state written (new)
};

__device__ __forceinline__ float shl_fill(float v, float fill) {  // lane+1's v; lane 63 takes fill
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float shr_fill(float v, float fill) {  // lane-1's v; lane 0 takes fill
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_of(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One plane of one component as a raw buffer: a wave-uniform descriptor plus a
// 32-bit lane byte offset per access; an offset at or past the plane's end
// (kOut) loads 0 and stores nothing.
constexpr uint32_t kOut = 0xffffffffu;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_of(const float *base, int64_t o, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base + o), 0, bytes, 0x00020000);
}
__device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t p, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p, off, 0, 0));
}
// Cache policy 2 is the non-temporal hint: the pass writes each output once
// and reads each coefficient once, so neither should displace from L2 the
// state planes that neighbouring tiles re-read (+4 % at 256^3).
constexpr int kNonTemporal = 2;
__device__ __forceinline__ float ld_once(__amdgpu_buffer_rsrc_t p, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(p, off, 0, kNonTemporal));
}
__device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t p, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), p, off, 0, kNonTemporal);
}

// Tile-edge values in the spare lanes: lane q holds row j0-1+q's old E and H
// at k0-1 with Ez at j+1 and i+1 and Ey, Ex at k0 (its new Hx, Hy), and its
// old Ex, Ey at k0+64.
struct Edge {
    float x, y, z, hx, hy, zj, zi, y0, x0;
    float ex, ey;
};

template <bool kMat, int R>
__global__ void __launch_bounds__(64 * kSGroup) __attribute__((amdgpu_waves_per_eu(kSWaves)))
k3_stream(wf_run3d r, Bufs b, int chunk, int ntk, int ntj, int n_tiles) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int l = (int)(threadIdx.x & 63u);
    // XCD-aware order: workgroup w runs on XCD w % 8; each XCD takes a
    // contiguous band of tiles (k fastest, then j, then i-chunk), so the rows
    // neighbouring tiles share meet in one L2
    const int per = (int)(gridDim.x >> 3);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // uniform: keeps tile math scalar
    const int t = ((int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3)) * kSGroup + wave;
    if (t >= n_tiles) return;  // whole wavefront
    const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
    const int k0 = tk * 64, j0 = tj * R, k = k0 + l;
    const int i0 = ti * chunk, i1 = min(nx, i0 + chunk);
    const int64_t si = (int64_t)ny * nz;
    const bool kin = k < nz;
    const bool kf = k <= nz - 2;
    const bool zinl = nz >= 4 && (r.flags & WF_RUN3D_SNAP_FROM_STATE) != 0;
    const bool kbox = zinl ? kin : (k >= 2 && k <= nz - 2);
    const bool ztile = zinl && (k0 == 0 || k0 + 64 >= nz);
    const bool kface = kin && (k == 0 || k == nz - 1);
    // rows q = 0 .. R+1 are j = j0-1+q; byte offsets inside a plane (the host
    // checks 4*ny*nz < 2^31), kOut outside the grid
    const int pb = 4 * ny * nz;
    uint32_t cq[R + 2];
    bool vq[R + 2];
#pragma unroll
    for (int q = 0; q < R + 2; ++q) {
        const int j = j0 - 1 + q;
        vq[q] = kin && j >= 0 && j < ny;
        cq[q] = vq[q] ? 4u * (uint32_t)(j * nz + k) : kOut;
    }
    const int jh = j0 - 1 + l;
    const bool hrow = l <= R && jh >= 0 && jh < ny;
    const bool lo_ok = hrow && k0 >= 1;
    const bool hi_ok = hrow && k0 + 64 < nz;
    const uint32_t c_lo = lo_ok ? 4u * (uint32_t)(jh * nz + k0 - 1) : kOut;
    const uint32_t c_lo1 = lo_ok ? c_lo + 4u : kOut;
    const uint32_t c_loj = lo_ok && jh + 1 < ny ? c_lo + 4u * (uint32_t)nz : kOut;
    const uint32_t c_hi = hi_ok ? 4u * (uint32_t)(jh * nz + k0 + 64) : kOut;

    int i = i0 > 0 ? i0 - 1 : 0;
    // old E of the current plane (carried from the previous iteration) and the
    // new Hy, Hz of the previous plane
    float ex[R + 1], ey[R + 1], ez[R + 1], exT, ezT;
    float hyp[R + 1], hzp[R + 1];
    {
        const int64_t o = (int64_t)i * si;
        const auto pex = plane_of(b.ex, o, pb), pey = plane_of(b.ey, o, pb), pez = plane_of(b.ez, o, pb);
#pragma unroll
        for (int q = 0; q <= R; ++q) {
            ex[q] = ld(pex, cq[q]);
            ey[q] = ld(pey, cq[q]);
            ez[q] = ld(pez, cq[q]);
            hyp[q] = hzp[q] = 0.f;
        }
        exT = ld(pex, cq[R + 1]);
        ezT = ld(pez, cq[R + 1]);
    }
    for (; i < i1; ++i) {
        const int64_t o = (int64_t)i * si;
        const bool ifar = i <= nx - 2;
        const int64_t o1 = ifar ? o + si : o;  // plane i+1 (unused on the last plane)
        const bool ib = i >= 2 && i <= nx - 2;
        // this plane's H, coefficients and tile-edge values, the next plane's E
        float ex1[R + 1], ey1[R + 1], ez1[R + 1], hx[R + 1], hy[R + 1], hz[R + 1], ca[R + 1], cb[R + 1];
        const auto pex = plane_of(b.ex, o, pb), pey = plane_of(b.ey, o, pb), pez = plane_of(b.ez, o, pb);
        const auto nex = plane_of(b.ex, o1, pb), ney = plane_of(b.ey, o1, pb), nez = plane_of(b.ez, o1, pb);
        const auto phx = plane_of(b.hx, o, pb), phy = plane_of(b.hy, o, pb), phz = plane_of(b.hz, o, pb);
        // loads issue in the order the arithmetic consumes them (the edge values
        // first), so each wait covers only what the next step needs
        Edge hc;
        hc.x = ld(pex, c_lo);
        hc.y = ld(pey, c_lo);
        hc.z = ld(pez, c_lo);
        hc.hx = ld(phx, c_lo);
        hc.hy = ld(phy, c_lo);
        hc.y0 = ld(pey, c_lo1);
        hc.x0 = ld(pex, c_lo1);
        hc.zj = ld(pez, c_loj);
        hc.zi = ld(nez, c_lo);
        hc.ex = ld(pex, c_hi);
        hc.ey = ld(pey, c_hi);
#pragma unroll
        for (int q = 0; q <= R; ++q) {
            ex1[q] = ld(nex, cq[q]);
            ey1[q] = ld(ney, cq[q]);
            ez1[q] = ld(nez, cq[q]);
            hx[q] = ld(phx, cq[q]);
            hy[q] = q > 0 ? ld(phy, cq[q]) : 0.f;  // row j0-1 needs no Hy
            hz[q] = ld(phz, cq[q]);
            ca[q] = cb[q] = 0.f;
            if (kMat && q > 0) {
                const int j = j0 - 1 + q;
                const uint32_t cc = ib && kbox && j >= 2 && j <= ny - 2 ? cq[q] : kOut;
                ca[q] = ld_once(plane_of(r.ca, o, pb), cc);
                cb[q] = ld_once(plane_of(r.cb, o, pb), cc);
            }
        }
        const float exT1 = ld(nex, cq[R + 1]), ezT1 = ld(nez, cq[R + 1]);
        // new Hx, Hy at (j0-1+l, k0-1) in the spare lanes (Faraday, fdtd3d.py:309-322)
        const float lhx = jh <= ny - 2 ? hc.hx + r.dh * (((hc.y0 - hc.y) / r.dz) - ((hc.zj - hc.z) / r.dy)) : hc.hx;
        const float lhy = ifar ? hc.hy + r.dh * (((hc.zi - hc.z) / r.dx) - ((hc.x0 - hc.x) / r.dz)) : hc.hy;
        // Faraday (fdtd3d.py:309-330): Hx on [:, :-1, :-1], Hy on [:-1, :, :-1], Hz on [:-1, :-1, :]
        float fhx[R + 1], fhy[R + 1], fhz[R + 1];
#pragma unroll
        for (int q = 0; q <= R; ++q) {
            const int j = j0 - 1 + q;
            const bool jf = j <= ny - 2;
            const float eyk = shl_fill(ey[q], lane_of(hc.ey, q));
            const float ezj = q < R ? ez[q + 1] : ezT;
            const float exj = q < R ? ex[q + 1] : exT;
            fhx[q] = jf && kf ? hx[q] + r.dh * (((eyk - ey[q]) / r.dz) - ((ezj - ez[q]) / r.dy)) : hx[q];
            fhz[q] = ifar && jf ? hz[q] + r.dh * (((exj - ex[q]) / r.dy) - ((ey1[q] - ey[q]) / r.dx)) : hz[q];
            if (q > 0) {
                const float exk = shl_fill(ex[q], lane_of(hc.ex, q));
                fhy[q] = ifar && kf ? hy[q] + r.dh * (((ez1[q] - ez[q]) / r.dx) - ((exk - ex[q]) / r.dz)) : hy[q];
            } else {
                fhy[q] = 0.f;
            }
        }
        // Mur z faces of Hx, Hy (boundaries.py:557-575): f[face] = P[in] + cz (f[in] - P[face])
        if (ztile) {
            const bool iz = i >= 1 && i <= nx - 2;
#pragma unroll
            for (int q = 0; q <= R; ++q) {
                const int j = j0 - 1 + q;
                const bool zf = iz && kface && j >= 1 && j <= ny - 2;
                // both shifts run on the whole wavefront before the select: a
                // DPP move under a divergent branch reads disabled lanes as absent
                {
                    const float fn = shl_fill(fhx[q], 0.f), fp = shr_fill(fhx[q], lane_of(lhx, q));
                    const float pn = shl_fill(hx[q], 0.f), pp = shr_fill(hx[q], lane_of(hc.hx, q));
                    const float fin = k == 0 ? fn : fp, pin = k == 0 ? pn : pp;
                    fhx[q] = zf ? pin + r.cmz * (fin - hx[q]) : fhx[q];
                }
                if (q > 0) {
                    const float fn = shl_fill(fhy[q], 0.f), fp = shr_fill(fhy[q], lane_of(lhy, q));
                    const float pn = shl_fill(hy[q], 0.f), pp = shr_fill(hy[q], lane_of(hc.hy, q));
                    const float fin = k == 0 ? fn : fp, pin = k == 0 ? pn : pp;
                    fhy[q] = zf ? pin + r.cmz * (fin - hy[q]) : fhy[q];
                }
            }
        }
        // new H of the core rows, and Ampere (fdtd3d.py:336-359) on the box
        const bool keep = i >= i0;
#pragma unroll
        for (int q = 1; q <= R; ++q) {
            const int j = j0 - 1 + q;
            const float hxm = shr_fill(fhx[q], lane_of(lhx, q));  // Hx(k-1)
            const float hym = shr_fill(fhy[q], lane_of(lhy, q));  // Hy(k-1)
            const bool box = ib && kbox && vq[q] && j >= 2 && j <= ny - 2;
            const float cxv = ((fhz[q] - fhz[q - 1]) / r.dy) - ((fhy[q] - hym) / r.dz);
            const float cyv = ((fhx[q] - hxm) / r.dz) - ((fhz[q] - hzp[q]) / r.dx);
            const float czv = ((fhy[q] - hyp[q]) / r.dx) - ((fhx[q] - fhx[q - 1]) / r.dy);
            float vx = ex[q], vy = ey[q];
            if (k >= 1) {
                vx = kMat ? ca[q] * ex[q] + cb[q] * cxv : ex[q] + r.de * cxv;
                vy = kMat ? ca[q] * ey[q] + cb[q] * cyv : ey[q] + r.de * cyv;
            }
            const float vz = kMat ? ca[q] * ez[q] + cb[q] * czv : ez[q] + r.de * czv;
            if (keep && vq[q]) {
                st(plane_of(b.nhx, o, pb), cq[q], fhx[q]);
                st(plane_of(b.nhy, o, pb), cq[q], fhy[q]);
                st(plane_of(b.nhz, o, pb), cq[q], fhz[q]);
                if (box) {
                    st(plane_of(b.nex, o, pb), cq[q], vx);
                    st(plane_of(b.ney, o, pb), cq[q], vy);
                    st(plane_of(b.nez, o, pb), cq[q], vz);
                }
            }
            hyp[q] = fhy[q];
            hzp[q] = fhz[q];
        }
#pragma unroll
        for (int q = 0; q <= R; ++q) {
            ex[q] = ex1[q];
            ey[q] = ey1[q];
            ez[q] = ez1[q];
        }
        exT = exT1;
        ezT = ezT1;
    }
}

// Shell of an axis of length n: the indices outside [2, n-2] ({0, 1, n-1} for
// n >= 4, {0} for n = 1), and the count of indices inside.
__device__ __forceinline__ int shell_count(int n) { return n == 1 ? 1 : 3; }
__device__ __forceinline__ int shell_index(int t, int n) { return t < 2 ? t : n - 1; }

// E on every cell the fused kernel leaves: the planes i in {0, 1, nx-1}, the
// rows j in {0, 1, ny-1} of the other planes, and the k-slab of the rest
// unless the fused pass applied the z faces;
// updated where the reference updates it (from the final new H), copied where
// it never does.
template <bool kMat>
__global__ void __launch_bounds__(256) k3_eshell(wf_run3d r, Bufs b) {
    const int nx = r.nx, ny = r.ny, nz = r.nz;
    const int sx = shell_count(nx), sy = shell_count(ny), sz = shell_count(nz);
    const int mx = nx - sx, my = ny - sy;
    const int64_t n1 = (int64_t)sx * ny * nz, n2 = (int64_t)mx * sy * nz;
    // the k-slab {0, 1, nz-1} of the remaining columns, unless the fused pass
    // applied the z faces itself and so updated every k there
    const bool zinl = nz >= 4 && (r.flags & WF_RUN3D_SNAP_FROM_STATE) != 0;
    const int64_t n3 = zinl ? 0 : (int64_t)mx * my * sz;
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
    if ((int64_t)r.nx * r.ny * r.nz > INT32_MAX) {  // cells are int32 in the source and detector tables
        wf::set_error("run3d: %dx%dx%d cells exceed the 2^31-cell index range of the source and detector tables",
                      r.nx, r.ny, r.nz);
        return WF_ERR_UNSUPPORTED;
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
            hipLaunchKernelGGL(k3_mur1<false>, dim3((unsigned)((m1 + 255) / 256)), dim3(256), 0, stream, r, Prev{});
            const int64_t m2 = 2 * (int64_t)r.nx * r.nz + 4 * zface_columns<false>(r.nx, r.ny, r.nz);
            hipLaunchKernelGGL(k3_mur2<false>, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, stream, r, Prev{});
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

// Resident wavefronts of the device the stream runs on, for sizing i-chunks.
static int resident_waves() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256 * 4 * kSWaves;
    if (!cached[dev]) {
        int n_cu = 0;
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
            n_cu = 256;
        cached[dev] = n_cu * 4 * kSWaves;
    }
    return cached[dev];
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
        if (4 * (int64_t)r.ny * r.nz > INT32_MAX) {
            wf::set_error("wf_fdtd3d_fused_step: an (ny, nz) plane of %d x %d cells exceeds 2^31 bytes", r.ny, r.nz);
            return WF_ERR_UNSUPPORTED;
        }
    }
    const int resident = resident_waves();
    for (int s = 0; s < n_steps; ++s) {
        const bool odd = s & 1;
        for (int bi = 0; bi < n_runs; ++bi) {
            const wf_run3d rd = view_of(runs[bi], odd);   // read state
            const wf_run3d wr = view_of(runs[bi], !odd);  // written state
            const int n = runs[bi].step + s;
            const bool from_state = (rd.flags & WF_RUN3D_SNAP_FROM_STATE) != 0;
            if (!from_state) {
                const Strips S(rd.nx, rd.ny, rd.nz);
                const unsigned snap_blocks = (unsigned)std::min<int64_t>((S.total + 255) / 256, 4096);
                hipLaunchKernelGGL(k3_snap, dim3(snap_blocks), dim3(256), 0, stream, rd);
            }
            if (rd.n_grp > 0) hipLaunchKernelGGL(k3_src, dim3(1), dim3(256), 0, stream, rd, n);
            Bufs bf{rd.ex, rd.ey, rd.ez, rd.hx, rd.hy, rd.hz, wr.ex, wr.ey, wr.ez, wr.hx, wr.hy, wr.hz};
            // i-chunks sized so the tiles are about one round of resident wavefronts
            const int ntk = (rd.nz + 63) / 64, ntj = (rd.ny + kSRows - 1) / kSRows;
            const int tiles = ntk * ntj;
            int n_chunks = std::max(1, std::min(std::max(1, rd.nx / 8), resident / std::max(1, tiles)));
            const int chunk = (rd.nx + n_chunks - 1) / n_chunks;
            n_chunks = (rd.nx + chunk - 1) / chunk;
            const int n_tiles = tiles * n_chunks;
            const dim3 gf(8u * (unsigned)(((n_tiles + kSGroup - 1) / kSGroup + 7) / 8)), bs(64 * kSGroup);
            if (rd.ca)
                hipLaunchKernelGGL((k3_stream<true, kSRows>), gf, bs, 0, stream, rd, bf, chunk, ntk, ntj, n_tiles);
            else
                hipLaunchKernelGGL((k3_stream<false, kSRows>), gf, bs, 0, stream, rd, bf, chunk, ntk, ntj, n_tiles);
            const Prev pv{rd.hx, rd.hy, rd.hz};
            const int64_t m1 = 4 * (int64_t)wr.ny * wr.nz + 2 * (int64_t)wr.nx * wr.nz;
            const unsigned g1 = (unsigned)((m1 + 255) / 256);
            if (from_state) {
                hipLaunchKernelGGL(k3_mur1<true>, dim3(g1), dim3(256), 0, stream, wr, pv);
                const int64_t m2 = 2 * (int64_t)wr.nx * wr.nz + 4 * zface_columns<true>(wr.nx, wr.ny, wr.nz);
                hipLaunchKernelGGL(k3_mur2<true>, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, stream, wr, pv);
            } else {
                hipLaunchKernelGGL(k3_mur1<false>, dim3(g1), dim3(256), 0, stream, wr, pv);
                const int64_t m2 = 2 * (int64_t)wr.nx * wr.nz + 4 * zface_columns<false>(wr.nx, wr.ny, wr.nz);
                hipLaunchKernelGGL(k3_mur2<false>, dim3((unsigned)((m2 + 255) / 256)), dim3(256), 0, stream, wr, pv);
            }
            const bool zinl = from_state && wr.nz >= 4;
            const int sx = wr.nx == 1 ? 1 : 3, sy = wr.ny == 1 ? 1 : 3, sz = wr.nz == 1 ? 1 : 3;
            const int64_t nsh = (int64_t)sx * wr.ny * wr.nz + (int64_t)(wr.nx - sx) * sy * wr.nz +
                                (zinl ? 0 : (int64_t)(wr.nx - sx) * (wr.ny - sy) * sz);
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
