/* Independent per-cell loop restatement of one 2D TM FDTD step (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: built by oracle/Makefile and loaded by tests to
 * pin oracle/fdtd2d_ref.py (the torch restatement). It is written from the
 * update equations as plain scalar loops, so a slicing mistake in the torch
 * restatement (an off-by-one range, a swapped operand, a wrong sign) shows up
 * as a mismatch. Built with -ffp-contract=off so every product and sum is
 * rounded separately, as eager torch does on the CPU.
 *
 * Semantics followed (reference file:line):
 *   snapshot before injection      src/core/fdtd2d.py:219-224, boundaries.py:237
 *   Faraday on [0,Nx-1)x[0,Ny-1)   src/core/fdtd2d.py:232-235
 *   Mur x-edges then y-edges       src/core/boundaries.py:259-274
 *   Ampere, free / material        src/core/fdtd2d.py:244-257
 *
 * Layout: row-major (i, j), index i*Ny + j, one grid (no batch).
 *
 * Division: the plain entry points divide by dx, dy (IEEE x / fp32(d), what
 * torch computes for the reference's `x / d` on the CPU). The *_rcp entry
 * points take the reciprocals fp32(1/dx), fp32(1/dy) in the dx, dy arguments
 * and multiply (x * fp32(1/d), what torch computes for the same expression
 * on a GPU).
 */
/* built with -fopenmp: wf_brute_run_f32 runs each phase's rows in parallel */
#include <stdlib.h>
#include <string.h>

#define IDX(i, j) ((size_t)(i) * (size_t)Ny + (size_t)(j))

#define DEFINE_STEP(NAME, T, QOP)                                                   \
void NAME(int Nx, int Ny, T *Ex, T *Ey, T *Hz, const T *Ca, const T *Cb,           \
          T Dh, T De, T dx, T dy, T cx, T cy,                                       \
          int n_inj, const int *inj_comp, const int *inj_cell, const T *inj_val)   \
{                                                                                   \
    size_t n = (size_t)Nx * (size_t)Ny;                                             \
    T *P = (T *)malloc(n * sizeof(T));                                              \
    T *Hn = (T *)malloc(n * sizeof(T));                                             \
    memcpy(P, Hz, n * sizeof(T));                 /* 1. snapshot */                 \
    for (int s = 0; s < n_inj; ++s) {             /* 2. soft sources, in order */   \
        T *f = inj_comp[s] == 0 ? Ex : (inj_comp[s] == 1 ? Ey : Hz);                \
        f[inj_cell[s]] = f[inj_cell[s]] + inj_val[s];                               \
    }                                                                               \
    memcpy(Hn, Hz, n * sizeof(T));                                                  \
    for (int i = 0; i < Nx - 1; ++i)              /* 3. Faraday */                  \
        for (int j = 0; j < Ny - 1; ++j) {                                          \
            T a = (Ex[IDX(i, j + 1)] - Ex[IDX(i, j)]) QOP dy;                       \
            T b = (Ey[IDX(i + 1, j)] - Ey[IDX(i, j)]) QOP dx;                       \
            T c = a - b;                                                            \
            T d = Dh * c;                                                           \
            Hn[IDX(i, j)] = Hz[IDX(i, j)] + d;                                      \
        }                                                                           \
    /* 4. Mur: left and right rows first (every j), then bottom/top columns */     \
    for (int j = 0; j < Ny; ++j) {                                                  \
        Hn[IDX(0, j)] = P[IDX(1, j)] + cx * (Hn[IDX(1, j)] - P[IDX(0, j)]);         \
    }                                                                               \
    for (int j = 0; j < Ny; ++j) {                                                  \
        Hn[IDX(Nx - 1, j)] = P[IDX(Nx - 2, j)]                                      \
                            + cx * (Hn[IDX(Nx - 2, j)] - P[IDX(Nx - 1, j)]);        \
    }                                                                               \
    for (int i = 0; i < Nx; ++i) {                                                  \
        Hn[IDX(i, 0)] = P[IDX(i, 1)] + cy * (Hn[IDX(i, 1)] - P[IDX(i, 0)]);         \
    }                                                                               \
    for (int i = 0; i < Nx; ++i) {                                                  \
        Hn[IDX(i, Ny - 1)] = P[IDX(i, Ny - 2)]                                      \
                            + cy * (Hn[IDX(i, Ny - 2)] - P[IDX(i, Ny - 1)]);        \
    }                                                                               \
    memcpy(Hz, Hn, n * sizeof(T));                                                  \
    /* 5. Ampere from the post-Mur Hz */                                           \
    for (int i = 0; i < Nx; ++i)                                                    \
        for (int j = 1; j < Ny; ++j) {                                              \
            T g = (Hz[IDX(i, j)] - Hz[IDX(i, j - 1)]) QOP dy;                       \
            if (Ca) {                                                               \
                T t1 = Ca[IDX(i, j)] * Ex[IDX(i, j)];                               \
                T t2 = Cb[IDX(i, j)] * g;                                           \
                Ex[IDX(i, j)] = t1 + t2;                                            \
            } else {                                                                \
                T t2 = De * g;                                                      \
                Ex[IDX(i, j)] = Ex[IDX(i, j)] + t2;                                 \
            }                                                                       \
        }                                                                           \
    for (int i = 1; i < Nx; ++i)                                                    \
        for (int j = 0; j < Ny; ++j) {                                              \
            T g = (Hz[IDX(i, j)] - Hz[IDX(i - 1, j)]) QOP dx;                       \
            if (Ca) {                                                               \
                T t1 = Ca[IDX(i, j)] * Ey[IDX(i, j)];                               \
                T t2 = Cb[IDX(i, j)] * g;                                           \
                Ey[IDX(i, j)] = t1 - t2;                                            \
            } else {                                                                \
                T t2 = De * g;                                                      \
                Ey[IDX(i, j)] = Ey[IDX(i, j)] - t2;                                 \
            }                                                                       \
        }                                                                           \
    free(P);                                                                        \
    free(Hn);                                                                       \
}

DEFINE_STEP(wf_brute_step_f32, float, /)
DEFINE_STEP(wf_brute_step_f64, double, /)
DEFINE_STEP(wf_brute_step_f32_rcp, float, *)

/* n_steps of the same step on one grid, for long runs at full size: the
 * snapshot and Faraday buffers are allocated once and every phase runs its
 * rows in parallel (OpenMP). Each phase writes every cell once from values
 * of the phase before, so the result does not depend on the thread count
 * and equals n_steps calls of wf_brute_step_f32. Sources: n_src
 * contributions applied in list order, contribution s adding
 * src_vals[s*n_steps + n] to its cell at step n (the caller forms
 * w[n]*amp as the reference does). Detectors: traces[d*n_steps + n] = Hz at
 * det_cell[d] after step n. */
static void run_f32(int Nx, int Ny, float *Ex, float *Ey, float *Hz, const float *Ca, const float *Cb,
                    float Dh, float De, float dx, float dy, float cx, float cy, int n_steps,
                    int n_src, const int *src_comp, const int *src_cell, const float *src_vals,
                    int n_det, const int *det_cell, float *traces, int rcp)
{
    size_t n = (size_t)Nx * (size_t)Ny;
    float *P = (float *)malloc(n * sizeof(float));
    float *Hn = (float *)malloc(n * sizeof(float));
    for (int st = 0; st < n_steps; ++st) {
        #pragma omp parallel for schedule(static)
        for (long c = 0; c < (long)n; ++c) P[c] = Hz[c];                  /* 1. snapshot */
        for (int s = 0; s < n_src; ++s) {                                    /* 2. sources, in order */
            float *f = src_comp[s] == 0 ? Ex : (src_comp[s] == 1 ? Ey : Hz);
            f[src_cell[s]] = f[src_cell[s]] + src_vals[(size_t)s * n_steps + st];
        }
        #pragma omp parallel for schedule(static)
        for (long c = 0; c < (long)n; ++c) Hn[c] = Hz[c];
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < Nx - 1; ++i)                                     /* 3. Faraday */
            for (int j = 0; j < Ny - 1; ++j) {
                float xa = Ex[IDX(i, j + 1)] - Ex[IDX(i, j)];
                float xb = Ey[IDX(i + 1, j)] - Ey[IDX(i, j)];
                float a = rcp ? xa * dy : xa / dy;
                float b = rcp ? xb * dx : xb / dx;
                float c = a - b;
                float d = Dh * c;
                Hn[IDX(i, j)] = Hz[IDX(i, j)] + d;
            }
        for (int j = 0; j < Ny; ++j)                                         /* 4. Mur: rows, then columns */
            Hn[IDX(0, j)] = P[IDX(1, j)] + cx * (Hn[IDX(1, j)] - P[IDX(0, j)]);
        for (int j = 0; j < Ny; ++j)
            Hn[IDX(Nx - 1, j)] = P[IDX(Nx - 2, j)] + cx * (Hn[IDX(Nx - 2, j)] - P[IDX(Nx - 1, j)]);
        for (int i = 0; i < Nx; ++i)
            Hn[IDX(i, 0)] = P[IDX(i, 1)] + cy * (Hn[IDX(i, 1)] - P[IDX(i, 0)]);
        for (int i = 0; i < Nx; ++i)
            Hn[IDX(i, Ny - 1)] = P[IDX(i, Ny - 2)] + cy * (Hn[IDX(i, Ny - 2)] - P[IDX(i, Ny - 1)]);
        #pragma omp parallel for schedule(static)
        for (long c = 0; c < (long)n; ++c) Hz[c] = Hn[c];
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < Nx; ++i)                                         /* 5. Ampere */
            for (int j = 1; j < Ny; ++j) {
                float xg = Hz[IDX(i, j)] - Hz[IDX(i, j - 1)];
                float g = rcp ? xg * dy : xg / dy;
                if (Ca) {
                    float t1 = Ca[IDX(i, j)] * Ex[IDX(i, j)];
                    float t2 = Cb[IDX(i, j)] * g;
                    Ex[IDX(i, j)] = t1 + t2;
                } else {
                    float t2 = De * g;
                    Ex[IDX(i, j)] = Ex[IDX(i, j)] + t2;
                }
            }
        #pragma omp parallel for schedule(static)
        for (int i = 1; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                float xg = Hz[IDX(i, j)] - Hz[IDX(i - 1, j)];
                float g = rcp ? xg * dx : xg / dx;
                if (Ca) {
                    float t1 = Ca[IDX(i, j)] * Ey[IDX(i, j)];
                    float t2 = Cb[IDX(i, j)] * g;
                    Ey[IDX(i, j)] = t1 - t2;
                } else {
                    float t2 = De * g;
                    Ey[IDX(i, j)] = Ey[IDX(i, j)] - t2;
                }
            }
        for (int d = 0; d < n_det; ++d) traces[(size_t)d * n_steps + st] = Hz[det_cell[d]];
    }
    free(P);
    free(Hn);
}

void wf_brute_run_f32(int Nx, int Ny, float *Ex, float *Ey, float *Hz, const float *Ca, const float *Cb,
                      float Dh, float De, float dx, float dy, float cx, float cy, int n_steps,
                      int n_src, const int *src_comp, const int *src_cell, const float *src_vals,
                      int n_det, const int *det_cell, float *traces)
{
    run_f32(Nx, Ny, Ex, Ey, Hz, Ca, Cb, Dh, De, dx, dy, cx, cy, n_steps, n_src, src_comp, src_cell, src_vals,
            n_det, det_cell, traces, 0);
}

/* The same run with reciprocal quotients: rdx = fp32(1/dx), rdy = fp32(1/dy). */
void wf_brute_run_f32_rcp(int Nx, int Ny, float *Ex, float *Ey, float *Hz, const float *Ca, const float *Cb,
                          float Dh, float De, float rdx, float rdy, float cx, float cy, int n_steps,
                          int n_src, const int *src_comp, const int *src_cell, const float *src_vals,
                          int n_det, const int *det_cell, float *traces)
{
    run_f32(Nx, Ny, Ex, Ey, Hz, Ca, Cb, Dh, De, rdx, rdy, cx, cy, n_steps, n_src, src_comp, src_cell, src_vals,
            n_det, det_cell, traces, 1);
}
