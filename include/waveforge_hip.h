/*
 * waveforge_hip.h -- C ABI of the MI355X (gfx950) 2D TM FDTD + DAS library
 * (libwaveforge_hip.so).
 *
 * The reference (shahzaibshazoo/waveforge) has no FFI: its boundary is the
 * Python object protocol of src/core (FDTD2D, MurABC, SourceCollection,
 * materials as Ca/Cb tensors) plus the DAS functions of its imaging
 * examples. Each entry point below replaces one eager-tensor sequence of that
 * protocol; the citation on each names the reference code it stands in for.
 *
 * Conventions
 *   - extern "C", plain pointers and sizes, no C++ types, no exceptions.
 *   - All pointers inside argument structs are DEVICE pointers unless a field
 *     says otherwise. The library allocates no device memory: callers pass
 *     every buffer (fields, Mur strips, tables, traces, workspaces).
 *   - Work is enqueued on the given stream and is asynchronous.
 *   - Return 0 on success, a negative WF_ERR_* code on failure; the message
 *     is available from wf_last_error() (thread-local).
 *   - Fields are fp32, row-major (i, j) with j fastest: cell = i*ny + j, the
 *     layout of a contiguous (Nx, Ny, 1) torch tensor (src/core/fields.py:156-168).
 */
#ifndef WAVEFORGE_HIP_H
#define WAVEFORGE_HIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_ABI_VERSION 3

#define WF_OK 0
#define WF_ERR_ARG (-1)
#define WF_ERR_HIP (-2)
#define WF_ERR_UNSUPPORTED (-3)

/* Bits of a run's device status word (wf_run2d.status, wf_run2d_f64.status,
 * wf_run3d.status). Kernels OR them in and never clear them; the caller
 * zeroes the word and reads it back through wf_status_check.
 *   WF_STATUS_WAVE_END     a step past the end of a source's waveform table:
 *                          that contribution was not applied (the reference
 *                          raises there, src/core/sources.py:882-887);
 *   WF_STATUS_TILE_GROUPS  a tile region of wf_fdtd2d_tb_advance held more
 *                          than 64 source groups (the extra ones were not
 *                          applied);
 *   WF_STATUS_TILE_DETS    a tile core of wf_fdtd2d_tb_advance held more than
 *                          64 detectors (the extra ones were not recorded). */
#define WF_STATUS_WAVE_END 1u
#define WF_STATUS_TILE_GROUPS 2u
#define WF_STATUS_TILE_DETS 4u

/* Field component codes of an injection (src/core/sources.py:49). The 2D
 * TM path uses the first three; the 3D path all six. */
#define WF_COMP_EX 0
#define WF_COMP_EY 1
#define WF_COMP_HZ 2
#define WF_COMP_EZ 3
#define WF_COMP_HX 4
#define WF_COMP_HY 5

/* One soft-source contribution: field[cell] += wave[n] * amp.
 * Contributions that hit the same (comp, cell) form a group and are applied in
 * array order, which is the SourceCollection order
 * (src/core/sources.py:1060-1071, point :882-887, line :1031-1035, plane
 * :715-736 with amp = 1). */
typedef struct wf_inj2d {
    const float *wave;   /* waveform table, wave_len entries */
    int32_t wave_len;
    int32_t comp;        /* WF_COMP_* */
    int32_t cell;        /* i*ny + j */
    float amp;
} wf_inj2d;

/* How a run divides by its cell sizes (wf_run2d.flags). The reference
 * divides float32 field differences by Python floats dx, dy
 * (src/core/fdtd2d.py:232-235, 250-251), and torch forms that quotient two
 * ways: on the CPU as IEEE x / fp32(d); on a GPU as x * fp32(1/d), the
 * reciprocal of the double d rounded once to fp32, then one fp32 product.
 * Without WF_RUN2D_RECIPROCAL the kernels reproduce the first, with it the
 * second (taking rdx, rdy from the descriptor). */
#define WF_RUN2D_RECIPROCAL 1

/* One simulation (one TX run / one batch member). The descriptor is read by
 * every entry point below; `step` is the global index of the next step. */
typedef struct wf_run2d {
    float *ex, *ey, *hz;          /* live state, (nx, ny) fp32 each */
    float *ex2, *ey2, *hz2;       /* second state buffer (temporally blocked path) */
    float *ex3, *ey3, *hz3;       /* third state buffer: NULL unless launches use parity 2 or 3 */
    const float *ca, *cb;         /* per-cell coefficients; NULL => free space */
    float *pstrip;                /* 4*(nx+ny) floats: Mur snapshot strips (per-step path) */
    const wf_inj2d *inj;          /* contributions sorted by group */
    const int32_t *grp_start;     /* n_grp+1 offsets into inj */
    const int32_t *det_cell;      /* n_det Hz detector cells (i*ny + j) */
    float *traces;                /* traces[d*trace_stride + (n - trace_n0)] */
    int64_t trace_stride;
    int32_t n_grp, n_inj;
    int32_t n_det, trace_n0, trace_len;
    int32_t nx, ny;
    float dh, de, dx, dy;         /* fp32(dt/MU0), fp32(dt/EPS0), fp32(dx), fp32(dy) */
    float cmx, cmy;               /* fp32 Mur coefficients (src/core/boundaries.py:177-178) */
    int32_t step;
    int32_t flags;                /* WF_RUN2D_* bits, else 0 */
    uint32_t *status;             /* NULL or device word: WF_STATUS_* bits (see above) */
    int32_t step_end;             /* 0, or the step index the run ends at: wf_fdtd2d_tb_advance takes no
                                     step n >= step_end (a launch that reaches it stops the run there, a
                                     launch that starts past it copies the run through) */
    float rdx, rdy;               /* fp32(1/dx), fp32(1/dy) of the double cell sizes: the multipliers
                                     of WF_RUN2D_RECIPROCAL */
    int32_t pad;
} wf_run2d;

int wf_abi_version(void);
const char *wf_last_error(void);

/* Wait for `stream`, read n_words device status words (WF_STATUS_* bits, as
 * the stepping kernels leave them) and return WF_OK if all are zero, else
 * WF_ERR_ARG with a message naming the first nonzero word and its
 * conditions. The words are not cleared. This is how a direct caller learns
 * of the errors the reference raises synchronously: stepping past a
 * waveform table (src/core/sources.py:882-887), and the blocked kernel's
 * per-tile source and detector limits. */
int wf_status_check(const uint32_t *status, int32_t n_words, hipStream_t stream);

/* Advance `n_steps` reference-exact leapfrog steps in place, starting at
 * run->step, one step per launch group (snapshot+inject, Faraday+Mur,
 * Ampere+detectors). Replaces one call of FDTD2D.step()
 * (src/core/fdtd2d.py:187-262) per step: MurABC.snapshot
 * (boundaries.py:230-237), SourceCollection.step (sources.py:1060-1071),
 * the Faraday slice update (fdtd2d.py:232-235), MurABC.apply
 * (boundaries.py:239-274), the Ampere slice updates (fdtd2d.py:244-257), and
 * the callers' detector gather (examples/brain_mimo_imaging.py:196-198).
 * `runs` is a HOST array of n_runs descriptors (batch members). A step past
 * a waveform table's end (the reference raises there, sources.py:882-887)
 * injects nothing and sets WF_STATUS_WAVE_END in run->status. Grids need 4 <= nx, ny and nx*ny < 2^31
 * (cells are int32 in the source and detector tables; WF_ERR_UNSUPPORTED
 * otherwise). */
int wf_fdtd2d_step(const wf_run2d *runs, int32_t n_runs, int32_t n_steps,
                   hipStream_t stream);

/* The same steps as wf_fdtd2d_step with no workspace from the caller and no
 * allocation: no counterpart in the reference. It is for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops, hosts
 * that forbid allocation). Each run's Mur snapshot strips (its pstrip
 * workspace, which is ignored here) live in a file-scope device array of the
 * library, 2^18 floats (1 MiB) per GPU, held until the process ends, so every
 * run needs 4*(nx+ny) <= 2^18 (nx + ny <= 65536; larger grids are rejected
 * with WF_ERR_ARG before any launch). The array's address is taken once per
 * device (hipGetSymbolAddress) on the first call with that device current;
 * call it with the stream's device current. All calls on one GPU share the
 * array: calls that overlap in time on one GPU (different streams) race; calls
 * on one stream do not. The strips need no clearing: every step writes them
 * before it reads them. Results are bit-identical to wf_fdtd2d_step. */
int wf_fdtd2d_step_noalloc(const wf_run2d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream);

/* max|ex|, max|ey|, max|hz| over n cells each -> out3[0..2] (device).
 * `partial` is a device workspace of at least 3*n_partial floats; n_partial
 * is the number of reduction blocks (1..4096). Replaces the three
 * tensor.abs().max() reductions of FDTD2D._check_stability
 * (src/core/fdtd2d.py:280-285); NaN propagates as in torch. */
int wf_fdtd2d_maxabs(const float *ex, const float *ey, const float *hz,
                     int64_t n, float *partial, int32_t n_partial,
                     float *out3, hipStream_t stream);

/* Temporally blocked advance of many runs in one launch per block of steps.
 * Each launch advances every run by `depth` steps (1..WF_TB_MAX_DEPTH),
 * reading one state buffer and writing another: parity 0 reads A (ex/ey/hz)
 * and writes B (ex2/ey2/hz2), 1 reads B and writes A, 2 reads A and writes C
 * (ex3/ey3/hz3), 3 reads C and writes B. An even number of launches
 * alternates 0, 1; an odd number (3 or more) runs 2, 3, 1 and then alternates,
 * so the state ends in A either way. Tiles of each run are listed in `tiles`
 * (device, 2 int32 per tile: run index, tile index within the run); a tile is
 * tile_i rows x tile_j columns, numbered row-major over ceil(nx/tile_i) x
 * ceil(ny/tile_j). tile_j = 64 - 2*depth stages one column per lane of each
 * 64-lane row (tile_i 8, 16, 32, 48 or 64); tile_j = 128 - 2*depth stages two
 * (tile_i 16, 32 or 48), which halves the share of recomputed halo columns.
 * Source, Mur and detector handling is exactly the per-step semantics above;
 * step indices are run->step + step_offset + level. A tile region may hold at
 * most 64 source groups and a tile core at most 64 detectors (more set
 * WF_STATUS_TILE_GROUPS / WF_STATUS_TILE_DETS in run->status and are not
 * applied); a step past a waveform table's end sets WF_STATUS_WAVE_END. If maxabs_out is non-NULL, the launch also folds
 * max|E|,|H| of its final level into maxabs_out[3*run + c] (as uint bit
 * patterns; the caller zeroes it first). fast_div selects how the launch
 * divides by dx, dy: WF_DIV_IEEE (IEEE x / fp32(d)); WF_DIV_FAST, the fast
 * quotient of wf_div_verify for every |x| < bits(div_span) (IEEE division
 * above it and for NaN), equal to IEEE division: callers enable it only for
 * divisors that wf_div_verify proved exact on [0, div_span), and div_lo must
 * then be 0; WF_DIV_RECIPROCAL, the product by each run's rdx, rdy (the
 * quotient of runs flagged WF_RUN2D_RECIPROCAL, so that the blocked path
 * and wf_fdtd2d_step give the same bits).
 * Replaces FDTD2D.run's Python loop over step() (src/core/fdtd2d.py:328-329)
 * and run_mimo's per-TX loop (examples/brain_mimo_imaging.py:176-201). */
#define WF_TB_MAX_DEPTH 8
#define WF_DIV_IEEE 0
#define WF_DIV_FAST 1
#define WF_DIV_RECIPROCAL 2
typedef struct wf_tb_args {
    const wf_run2d *runs;      /* device array of n_runs descriptors */
    const int32_t *tiles;      /* device, 2*n_tiles */
    int32_t n_runs, n_tiles;
    int32_t tile_i, tile_j;    /* core tile shape: tile_i from wf_tb_tile_shapes, tile_j = 64*cpl - 2*depth
                                  with cpl = 1 or 2 columns per lane */
    int32_t depth, parity;     /* parity 0..3: which state buffers the launch reads and writes */
    int32_t step_offset;
    int32_t has_materials;     /* 1 if every run has ca/cb, 0 if none has */
    uint32_t *maxabs_out;      /* NULL or device 3*n_runs words */
    float rdx, rdy;            /* WF_DIV_FAST: fp32 reciprocals 1/fp32(dx), 1/fp32(dy) (correctly rounded) */
    uint32_t div_lo_x, div_span_x;
    uint32_t div_lo_y, div_span_y;
    int32_t fast_div;          /* WF_DIV_*: WF_DIV_FAST needs all runs to share dx, dy */
    int32_t pad;
    const int32_t *stop;       /* NULL or device n_runs flags: a run whose flag is set is
                                  copied through unchanged (it failed a stability check) */
    int64_t hist_stride;       /* 0, or floats between consecutive state slots: the launch
                                  then writes the state after each of its `depth` steps s
                                  to the written state + (s-1)*hist_stride instead of only
                                  the last one (materials, IEEE division, tile_i 32 or 48,
                                  no maxabs_out or stop); the adjoint's recompute of a
                                  segment's state history */
    const int32_t *tile_lists; /* NULL, or device, 4 int32 per tile in the order of `tiles`:
                                  [grp_lo, grp_hi) into grp_list and [det_lo, det_hi) into
                                  det_list. grp_list[grp_lo..grp_hi) must be exactly the run's
                                  source groups whose cell lies in the tile's region (the core
                                  plus `depth` cells on every side), det_list[det_lo..det_hi)
                                  exactly its detectors in the core. With them a tile reads its
                                  own lists instead of testing every group and detector of its
                                  run; NULL keeps that scan. */
    const int32_t *grp_list;   /* group indices into the run's grp_start */
    const int32_t *det_list;   /* detector indices into the run's det_cell */
} wf_tb_args;

int wf_fdtd2d_tb_advance(const wf_tb_args *args, hipStream_t stream);

/* fp64 counterparts of wf_inj2d / wf_run2d for the double-precision check
 * path: the same step in fp64 arithmetic, as the reference runs when its grid
 * dtype is float64 (src/core/fields.py:156-168 allocates in the grid dtype;
 * the coefficients are then Python doubles, src/core/fdtd2d.py:158-159, and
 * MurABC's, boundaries.py:177-178). It serves as the FP64 reference run that
 * per-result errors of the fp32 path are stated against, and as the forward
 * of central-difference gradient checks. Same layout and ownership rules as
 * the fp32 descriptors; no temporally blocked variant. */
typedef struct wf_inj2d_f64 {
    const double *wave;
    int32_t wave_len;
    int32_t comp;
    int32_t cell;
    int32_t pad;
    double amp;
} wf_inj2d_f64;

typedef struct wf_run2d_f64 {
    double *ex, *ey, *hz;
    const double *ca, *cb;        /* NULL => free space */
    double *pstrip;               /* 4*(nx+ny) doubles */
    const wf_inj2d_f64 *inj;
    const int32_t *grp_start;
    const int32_t *det_cell;
    double *traces;
    int64_t trace_stride;
    int32_t n_grp, n_inj;
    int32_t n_det, trace_n0, trace_len;
    int32_t nx, ny;
    int32_t step;
    double dh, de, dx, dy;        /* dt/MU0, dt/EPS0, dx, dy in fp64 */
    double cmx, cmy;
    uint32_t *status;             /* NULL or device word: WF_STATUS_* bits */
} wf_run2d_f64;

typedef struct wf_f64_args {
    const wf_run2d_f64 *runs;     /* device array of n_runs descriptors */
    int32_t n_runs;
    int32_t max_nx, max_ny;       /* bounds of every run's nx, ny */
    int32_t n_steps;
    int32_t step_offset;          /* step index = runs[r].step + step_offset + s */
    int32_t pad;
} wf_f64_args;

/* Advance every run `n_steps` fp64 steps in place, three launches per step
 * covering all runs (the per-step semantics of wf_fdtd2d_step). Replaces
 * FDTD2D.step() on a float64 grid (src/core/fdtd2d.py:187-266). */
int wf_fdtd2d_step_f64(const wf_f64_args *args, hipStream_t stream);

/* wf_fdtd2d_maxabs for float64 fields: max|ex|, max|ey|, max|hz| -> out3
 * (device, 3 doubles); `partial` holds at least 3*n_partial doubles. The
 * stability check of FDTD2D on a float64 grid (src/core/fdtd2d.py:280-285). */
int wf_fdtd2d_maxabs_f64(const double *ex, const double *ey, const double *hz, int64_t n, double *partial,
                         int32_t n_partial, double *out3, hipStream_t stream);

/* One 3D simulation (FDTD3D). Fields are (nx, ny, nz) fp32, row-major with
 * k fastest: cell = (i*ny + j)*nz + k, the layout of a contiguous (Nx, Ny, Nz)
 * torch tensor (src/core/fields.py:156-168). nz = 1 is the reference's
 * 2D-equivalent mode (no z faces). Sources use wf_inj2d with any WF_COMP_*
 * code and 3D cell indices; detectors record component det_comp[d] at
 * det_cell[d] after every step. */
typedef struct wf_run3d {
    float *ex, *ey, *ez, *hx, *hy, *hz;
    float *ex2, *ey2, *ez2, *hx2, *hy2, *hz2;  /* second state (fused path); unused by wf_fdtd3d_step */
    const float *ca, *cb;         /* per-cell coefficients; NULL => free space */
    float *pstrip;                /* wf_fdtd3d_strip_floats(nx, ny, nz) floats: Mur snapshot planes */
    const wf_inj2d *inj;
    const int32_t *grp_start;
    const int32_t *det_cell;
    const int32_t *det_comp;
    float *traces;                /* traces[d*trace_stride + (n - trace_n0)] */
    int64_t trace_stride;
    int32_t n_grp, n_inj;
    int32_t n_det, trace_n0, trace_len;
    int32_t nx, ny, nz;
    float dh, de, dx, dy, dz;     /* fp32(dt/MU0), fp32(dt/EPS0), fp32 cell sizes */
    float cmx, cmy, cmz;          /* fp32 Mur coefficients (boundaries.py:468-470) */
    int32_t step;
    int32_t flags;                /* WF_RUN3D_* bits, else 0 */
    uint32_t *status;             /* NULL or device word: WF_STATUS_* bits */
} wf_run3d;

/* wf_run3d.flags bit, honoured by wf_fdtd3d_fused_step only: the Mur snapshot
 * (boundaries.py:495-504, taken before the sources are added) equals the read
 * state on every Mur plane, because no H-component source cell lies on a
 * plane at index 0, 1, n-2 or n-1 across either tangential axis of its
 * component. The step then takes no snapshot copy: the Mur faces read the
 * read state, and the z faces of Hx, Hy are applied inside the streaming
 * pass. Setting it when a source does lie on such a plane gives wrong faces. */
#define WF_RUN3D_SNAP_FROM_STATE 1

/* Floats of the Mur snapshot workspace of one 3D run: the planes at index
 * 0, 1, n-2, n-1 across each tangential axis of each H component,
 * 8*(ny*nz + nx*nz) + (nz > 1 ? 8*nx*ny : 0). */
int64_t wf_fdtd3d_strip_floats(int32_t nx, int32_t ny, int32_t nz);

/* Advance `n_steps` reference-exact 3D steps in place, starting at
 * run->step: MurABC3D.snapshot (boundaries.py:495-504), the sources, the
 * Faraday slice updates (fdtd3d.py:309-330), MurABC3D.apply
 * (boundaries.py:506-575), the Ampere slice updates (fdtd3d.py:336-359) and
 * the detectors; replaces one call of FDTD3D.step() (fdtd3d.py:262-373) per
 * step. `runs` is a HOST array of n_runs descriptors. */
int wf_fdtd3d_step(const wf_run3d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream);

/* The same steps with one fused pass per step: step s reads state A (ex..hz)
 * and writes state B (ex2..hz2) when s is even, the reverse when odd, so
 * `n_steps` must be even and the result lands in state A. Per step: the Mur
 * snapshot planes (skipped under WF_RUN3D_SNAP_FROM_STATE) and the sources on
 * the read state; one launch in which each wavefront streams a 4-row x 64-lane
 * (j, k) tile along i, computing Faraday everywhere, the z faces of Hx, Hy
 * away from the x and y faces (under the flag), and Ampere on the cells whose
 * stencil touches no face still to be written; then the Mur faces left, the
 * remaining E cells and the detectors. Results equal wf_fdtd3d_step's bit for
 * bit. Requires 4*ny*nz < 2^31 (WF_ERR_UNSUPPORTED otherwise). */
int wf_fdtd3d_fused_step(const wf_run3d *runs, int32_t n_runs, int32_t n_steps, hipStream_t stream);

/* Stability judgement on the device, so the stepping stream never waits for
 * the host: for each run r, if any of the three maxima maxabs[3r..3r+2]
 * (uint bit patterns of max|Ex|, max|Ey|, max|Hz|, as folded by
 * wf_fdtd2d_tb_advance) exceeds `threshold`, set stop[r] = 1; with all != 0,
 * one failing run sets every flag (a batched FieldSet fails as a whole).
 * Replaces the `val > threshold` test of FDTD2D._check_stability
 * (src/core/fdtd2d.py:285-295); NaN maxima never exceed, as in the reference. */
int wf_fdtd2d_judge(const uint32_t *maxabs, int32_t n_runs, double threshold, int32_t all, int32_t *stop,
                    hipStream_t stream);

/* runs[r].step += delta for the n_runs device descriptors. A stepping loop
 * captured in a HIP graph ends with this kernel, so each replay advances the
 * step index its sources and detectors use. No reference counterpart: it
 * serves the graph-captured form of the run loop (src/core/fdtd2d.py:328-329). */
int wf_runs_add_step(wf_run2d *runs, int32_t n_runs, int32_t delta, hipStream_t stream);

/* Supported tile shapes of wf_tb_args, written into out as (tile_i, cpl)
 * pairs (host, 2*count int32; cpl = columns per lane, so tile_j =
 * 64*cpl - 2*depth); returns count (<= cap). One column per lane: 8, 16,
 * 32, 48 and 64 rows; two columns per lane: 16, 32 and 48 rows. Every shape
 * runs 8-wavefront workgroups (the tile's rows plus 2*depth halo rows dealt
 * round-robin over the wavefronts). */
int wf_tb_tile_shapes(int32_t *out, int32_t cap);

/* Exhaustive device check that the fast quotient equals IEEE x/d bit for bit
 * for every float x with lo <= bits(|x|) < lo + span, both signs. With
 * d = m*2^e (normal, -95 <= e <= 31) and s = 2^(31-e) the fast quotient is
 * q0 = x*r, t = fma(q0, d*s, -(x*s)), q = fma(-t, r/s, q0): the residual
 * taken at scale s stays exact down to the smallest subnormal x.
 * Adds the number of mismatches to *bad (device uint32, caller zeroes).
 * No reference counterpart: it licenses the fast division of
 * wf_fdtd2d_tb_advance, which replaces the reference's tensor / dx. */
int wf_div_verify(float d, float r, uint32_t lo, uint32_t span, uint32_t *bad, hipStream_t stream);

/* One discrete adjoint step of the leapfrog update for n_runs runs that share
 * the grid and the coefficients (one phantom, several TX). Given the adjoint
 * state of u_{n+1} in (lex, ley, lhz), it adds dL/dd_n at the receivers, applies
 * the transposed stages of step n in reverse order (Ampere, Mur top/bottom/
 * right/left, Faraday, snapshot) leaving the adjoint of u_n in (lex, ley, lhz),
 * and accumulates dL/dCa and dL/dCb (fp64) per run. ex_prev/ey_prev are the
 * fields before step n, hz_next the Hz after it. Sources are affine and drop
 * out. With n_steps > 1 the call runs steps step, step-1, ..., step-n_steps+1
 * in turn, the state pointers moving back by hist_stride floats per step (a
 * segment of a state history laid out slot after slot). Two consecutive
 * steps share one launch when the receivers' records can be folded in (no
 * receivers, or rx_first/rx_order given) unless flags has WF_ADJ_ONE_STEP;
 * otherwise each step is one launch. Every launch reads the adjoint state
 * from one set and writes the other (lex2/ley2/lhz2 are the second set,
 * overwritten); the state is back in (lex, ley, lhz) when the call returns.
 * All three launch forms give the same bits. No reference code counterpart: the
 * reference's differentiable-FDTD plan is design prose
 * (docs/architecture/DIFFERENTIABLE_PHYSICS.md:33); the forward it
 * differentiates is FDTD2D.step (src/core/fdtd2d.py:187-266). */
#define WF_ADJ_TWO_PASS 1
#define WF_ADJ_ONE_STEP 2  /* one launch per step even where two steps could share one */
#define WF_ADJ_FAST_DIV 4  /* two-step launches divide by dx, dy with the verified fast quotient */

typedef struct wf_adj_args {
    const float *ca, *cb;            /* (nx, ny), shared by all runs */
    const float *ex_prev, *ey_prev;  /* run r at r*state_stride */
    const float *hz_next;
    float *lex, *ley, *lhz;          /* adjoint state, run r at r*run_stride */
    float *lex2, *ley2, *lhz2, *lp;  /* scratch, same layout (lp is no longer written; kept for the layout) */
    double *gca, *gcb;               /* gradient accumulators, run r at r*run_stride */
    const float *dl_dd;              /* dL/d trace[r][k][n] at r*dd_run_stride + k*dd_rx_stride + n */
    const int32_t *rx_cell;          /* n_rx receiver cells */
    int64_t run_stride, state_stride, dd_run_stride, dd_rx_stride;
    int32_t nx, ny, n_runs, n_rx, step;
    float dh, dx, dy, cmx, cmy;
    int32_t n_steps;                 /* adjoint steps in this call; 0 or 1: one */
    int64_t hist_stride;             /* floats between the state slots of consecutive steps */
    const int32_t *rx_first;         /* NULL, or (nx, ny): -1, or the position in rx_order of the
                                        first receiver at that cell */
    const int32_t *rx_order;         /* NULL, or the n_rx receiver indices ordered by cell (receivers
                                        of one cell in index order); with both, a cell's residuals
                                        are added in that order and all steps of a call after the
                                        first record inside the previous step's last kernel */
    int32_t flags;                   /* WF_ADJ_TWO_PASS: two launches per step (k_adj_e, k_adj_h)
                                        instead of one; WF_ADJ_ONE_STEP: one launch per step;
                                        WF_ADJ_FAST_DIV: the fields below hold verified divisions.
                                        Every form gives the same bits */
    int32_t pad2;
    float div_rx, div_ry;            /* with WF_ADJ_FAST_DIV: RN(1/dx), RN(1/dy) */
    uint32_t div_span_x, div_span_y; /* and the |x| bit ranges [0, span) on which wf_div_verify
                                        found the fast quotient by dx, dy equal to IEEE x/d */
} wf_adj_args;

int wf_adjoint_step(const wf_adj_args *args, hipStream_t stream);

/* Delay-and-sum backprojection of one trace set onto a pixel grid.
 * Replaces das_backprojection of examples/brain_mimo_imaging.py:211-266 and
 * its variants (time gate: examples/09_brain_clot_dataset_sample.py:176-178;
 * other medium speed: examples/10_breast_tumor_mimo.py:148-165; zero instead
 * of clip outside [0, T): src/datasets/brain/antenna.py:333-336). fp64
 * geometry and accumulation in the reference's tx-major, rx-minor order,
 * round-half-even sample index, image = sum^2. */
typedef struct wf_das2d_args {
    const float *traces;       /* (n_tx, n_rx, n_t) fp32, contiguous */
    const double *ant_x;       /* n_ant antenna x positions in metres */
    const double *ant_y;
    double *image;             /* (nx_img, ny_img) fp64 output */
    int32_t n_tx, n_rx, n_t;
    int32_t nx_img, ny_img;
    int32_t gate_start;        /* samples with index < gate_start read as 0 */
    int32_t clip;              /* 1: clip index to [0, n_t-1]; 0: drop outside */
    int32_t pad;
    double pitch_x, pitch_y;   /* pixel pitch in metres */
    double v_medium, dt;
} wf_das2d_args;

int wf_das2d(const wf_das2d_args *args, hipStream_t stream);

/* wf_das2d for n_images images at once: `args` is a HOST array of n_images
 * descriptors, each validated as wf_das2d validates it (images and trace sets
 * may differ in size). Up to 32 images share one launch. Each image equals
 * what wf_das2d gives for its descriptor, bit for bit. Replaces the
 * reference's one das_backprojection call per phantom
 * (examples/brain_mimo_imaging.py:211-266) when many images are formed at
 * once (all phantoms of a batch, all requests a serving round retires). */
int wf_das2d_batch(const wf_das2d_args *args, int32_t n_images, hipStream_t stream);

/* Envelope delay-and-sum image with layered-medium travel times: replaces
 * AntennaRing.compute_das_image (src/datasets/brain/antenna.py:262-333), the
 * DAS the dataset generator runs on background-subtracted signals
 * (src/datasets/generator.py:356-370). Stages (bitmask `stages`):
 *   1  env[t, n] = |analytic signal of traces[t, :]|, the analytic signal of
 *      SciPy's hilbert (antenna.py:294): real part x, imaginary part the
 *      circular convolution of x with hilbert_k (n_t fp64 values, host
 *      supplied), evaluated in fp64, stored fp32;
 *   2  tt[a, iy, ix] = one-way travel time in samples from antenna a to
 *      pixel (px[ix], py[iy]) (fp64, the reference's expression order, stored
 *      fp32): through concentric layers of radius layer_r (mm, outermost first)
 *      and permittivity layer_eps around (cx, cy) (antenna.py:218-260), or in
 *      free space when n_layers is 0 (antenna.py:314-321); speed 3e8 m/s;
 *   4  image[iy, ix] = (sum over tx-major, rx-minor pairs of
 *      env[tx, rx][rint(tt[tx] + tt[rx])], indices outside [0, n_t) adding
 *      nothing, accumulated in fp64)^2, stored fp32 (antenna.py:324-333).
 * Coordinates in mm, dt in seconds. Stages may run separately (e.g. stage 2
 * once per geometry, stages 1|4 per signal set). */
typedef struct wf_das_env_args {
    const float *traces;          /* (n_ant, n_ant, n_t) fp32 */
    const double *hilbert_k;      /* n_t */
    float *env;                   /* (n_ant, n_ant, n_t) fp32 */
    float *tt;                    /* (n_ant, img_h, img_w) fp32 */
    float *image;                 /* (img_h, img_w) fp32 */
    const double *ant_x, *ant_y;  /* n_ant, mm */
    const double *px, *py;        /* img_w, img_h pixel coordinates, mm */
    const double *layer_r, *layer_eps;
    double cx, cy;                /* layer centre, mm */
    double dt;
    int32_t n_ant, n_t, img_w, img_h, n_layers, stages;
} wf_das_env_args;

int wf_das_envelope(const wf_das_env_args *args, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* WAVEFORGE_HIP_H */
