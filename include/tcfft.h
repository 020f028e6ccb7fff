/*
 * tcfft.h -- C ABI of the MI355X (gfx950) half-precision FFT library
 * (libtcfft.so).  Plain pointers, sizes and status codes; no HIP or torch
 * types.  `stream` arguments are hipStream_t values passed as void*
 * (NULL = the null stream, as the reference launches on).
 *
 * Data: interleaved FP16 complex (re, im halves), signal b of length n at
 * element offset n*b; 2D images row-major [batch][nx][ny] with ny contiguous,
 * exactly the buffers of tcfft_half.h / tcfft_half_2d.h.  Forward is
 * exp(-2*pi*i*nk/N) (FFTW_FORWARD, tcfft_half.cu:40-44), unnormalised.
 *
 * The C++ drop-in headers tcfft_half.h and tcfft_half_2d.h are thin inline
 * shims over these entry points.
 */
#ifndef TCFFT_H
#define TCFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcfft_plan_s *tcfft_plan;

enum tcfft_status {
    TCFFT_SUCCESS = 0,
    TCFFT_INVALID_SIZE = 1,    /* n / nx / ny not a supported power of two */
    TCFFT_INVALID_BATCH = 2,   /* batch < 1 or total elements >= 2^31 */
    TCFFT_INVALID_POINTER = 3, /* NULL or misaligned (< 16 B) buffer, NULL out-parameter */
    TCFFT_INVALID_PLAN = 4,    /* NULL plan, or a plan of the other kind */
    TCFFT_INVALID_VALUE = 5,   /* bad direction, axis, capacity ... */
    TCFFT_HIP_ERROR = 6,       /* a HIP runtime call failed (see tcfft_last_hip_error) */
    TCFFT_ALLOC_FAILED = 7
};

enum tcfft_direction { TCFFT_FORWARD = -1, TCFFT_INVERSE = +1 };

/* Library version, e.g. 100 for 0.1.0. */
int tcfft_version(void);

/* Human-readable name of a status code (static storage). */
const char *tcfft_status_string(int status);

/* Last HIP error code seen by the library on this thread (0 if none). */
int tcfft_last_hip_error(void);

/* Replaces tcfftCreate(tcfftHandle*, int n, int n_batch), tcfft_half.h:21 /
 * tcfft_half.cu:1235-1345.  n is a power of two, 256 <= n <= 2^27; the
 * reference's table covers only 10 of these sizes (tcfft_half.cu:1250-1330).
 * Uploads FP64-generated twiddle tables to the current HIP device. */
int tcfft_plan_create_1d(tcfft_plan *plan, int n, int n_batch);

/* Replaces tcfftCreate(tcfftHandle*, int nx, int ny, int n_batch),
 * tcfft_half_2d.h:21 / tcfft_half_2d.cu:1151-1221.  256 <= ny <= 2^16 and
 * 2 <= nx <= 2^16 powers of two (the reference: nx in {256,512},
 * ny in {256,512,1024}). */
int tcfft_plan_create_2d(tcfft_plan *plan, int nx, int ny, int n_batch);

/* Radix split of a 1D plan.  first_log2 = log2 of the first (contiguous)
 * pass's length when it does not cover the whole line: 8..12 (a row pass of
 * 256..4096 points) or 16 (the 65536-point single-pass kernel, the default);
 * the remaining bits go to one or two strided passes.  The radix list
 * reported by tcfft_plan_radices follows the split.  No reference
 * counterpart (tcfft_half.cu:1250-1330 has one fixed split per size). */
int tcfft_plan_create_1d_split(tcfft_plan *plan, int n, int n_batch, int first_log2);
/* Builds the candidate splits (first pass of 65536 points or the whole line,
 * and 4096 points), times each on a scratch buffer of the plan's size (the
 * call allocates and frees it, on the current device's null stream) and
 * returns the fastest plan; *first_log2 and *best_ms report the choice. */
int tcfft_plan_create_1d_tuned(tcfft_plan *plan, int n, int n_batch, int *first_log2, float *best_ms);
int tcfft_query_1d_split(int n, int n_batch, int first_log2, int *radices, int capacity, int *n_radices,
                         int *n_passes);

/* The reference never frees a plan (tcfft_half.cu:1331-1344 leaks); this does. */
int tcfft_plan_destroy(tcfft_plan plan);

/* Digit-reversal radix list the plan's passes consume along `axis`
 * (1D: axis 0; 2D: axis 0 = x (columns, length nx), axis 1 = y (rows, length
 * ny)).  Replaces reading tcfftHandle::radices / n_radices
 * (tcfft_doit_half.cpp:38) and radices_x/_y (tcfft_doit_half_2d.cpp:40-41). */
int tcfft_plan_radices(tcfft_plan plan, int axis, int *radices, int capacity, int *n_radices);

/* Host-only query of the 1D plan structure (no device access): the radix
 * list and the number of passes tcfft_plan_create_1d would build. */
int tcfft_query_1d(int n, int n_batch, int *radices, int capacity, int *n_radices, int *n_passes);
int tcfft_query_2d(int nx, int ny, int n_batch, int *radices_x, int *n_radices_x, int *radices_y,
                   int *n_radices_y, int *n_passes, int capacity);

/* Replaces tcfftExec(tcfftHandle, half *data), tcfft_half.h:20 /
 * tcfft_half.cu:1177-1233 (and the 2D one, tcfft_half_2d.cu:1128-1149):
 * in place on a device buffer whose input is digit-reversed by the plan's
 * radix list (per axis for 2D); output in natural order.  direction is
 * TCFFT_FORWARD or TCFFT_INVERSE (unnormalised adjoint).  Asynchronous on
 * `stream`. */
int tcfft_exec(tcfft_plan plan, void *data, int direction, void *stream);

/* Natural-order input and output; the result equals tcfft_exec on the input
 * permuted by the caller-side gen_rev (tcfft_doit_half.cpp:40-46), bit for
 * bit.  The digit reversal runs in the first pass's load for lines (1D, or
 * 2D rows) of up to 65536 points, else as one extra tiled reversal pass.  in == out is allowed; in and out must not otherwise
 * overlap.  The 2D in-place case and the 1D in-place case above 65536
 * points stage the input in a device workspace of the buffer's size that such
 * plans allocate at creation (2D plans, 1D plans whose first pass does not
 * read natural order), so concurrent in-place calls on one plan from
 * different streams race on it.  No call allocates, frees or memsets: every
 * case can be captured into a hipGraph. */
int tcfft_exec_natural(tcfft_plan plan, const void *in, void *out, int direction, void *stream);

/* tcfft_exec_natural that allocates nothing and uses no plan workspace (no
 * counterpart in the reference): for callers that cannot allocate device
 * memory -- code under hipGraph capture, serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, kernels and results (bit for bit) as
 * tcfft_exec_natural.  Its in-place cases (2D plans; 1D plans whose first
 * pass does not read natural order, e.g. n > 65536) stage the input in a
 * fixed device array of TCFFT_NOALLOC_SCRATCH_BYTES per GPU that the library
 * holds until the process ends; a call that would stage more returns
 * TCFFT_INVALID_SIZE.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (several streams or host threads) race on it and must
 * be serialised by the caller.  No allocation, free or memset. */
#define TCFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tcfft_exec_natural_noalloc(tcfft_plan plan, const void *in, void *out, int direction, void *stream);

/* The caller-side staging of tcfftExec on the device: for a 1D plan of n
 * points, out[b*n + p] = in[b*n + rev[p]] for n_lines lines (rev = the
 * plan's gen_rev order, tcfft_doit_half.cpp:6-29,40-46), as one LDS-tiled
 * pass at copy speed.  4-byte elements (fp16 complex), out of place (no
 * overlap), 16-byte aligned; n_lines is independent of the plan's batch.
 * Allocates nothing; graph-capturable. */
int tcfft_permute_to_dropin(tcfft_plan plan, const void *in, void *out, int n_lines, void *stream);

/* Split precision (no counterpart in the reference; config "FP16 hi+lo of
 * FP32 inputs"): FP32 complex input and output (interleaved float re/im,
 * natural order, the plan's 1D or 2D layout), every radix-16 stage as FP16
 * hi/lo products on MFMA with FP32 accumulation, radix 2..8 steps in FP32;
 * ~1e-7 relative error.  A plan prepared by tcfft_plan_prepare_split takes
 * every size it was created for, in place or out of place; an unprepared
 * plan only 1D n = 65536 out of place (else TCFFT_INVALID_SIZE, or
 * TCFFT_INVALID_POINTER for in == out).  Partially overlapping buffers are
 * rejected.  Real-input plans: TCFFT_INVALID_PLAN. */
int tcfft_exec_split(tcfft_plan plan, const void *in, void *out, int direction, void *stream);

/* Prepares a complex plan for tcfft_exec_split at its size: builds the
 * pass list (1D n = 65536 as two 256 x 256 passes; other shapes as passes
 * of 16..256 rows, 2..4 per axis), uploads their twiddle tables and
 * allocates a workspace of 8 bytes per element, so that tcfft_exec_split
 * allocates nothing (graph-capturable) and runs in place.  One workspace per
 * plan: split calls on one plan must not overlap (out-of-place calls on a 1D
 * plan of at most 65536 points do not use it and may), and this call must not
 * run concurrently with any other call on the plan.  Idempotent.  No
 * counterpart in the reference. */
int tcfft_plan_prepare_split(tcfft_plan plan);

/* The split-precision pass list tcfft_plan_prepare_split builds for a shape
 * (dims 1: n_or_nx = n, ny ignored; dims 2: nx x ny), computed on the host
 * without a device.  Pass i fills desc[5i .. 5i+4] = {R, log2 Q, log2 Qlo,
 * log2 M or 0, units}: each signal seen as R rows x Q columns, the R-point
 * DFT down every column, output k of column q = q_hi*Qlo + q_lo times
 * W_M^{q_hi k} (0: no twiddle) stored at q_hi*R*Qlo + k*Qlo + q_lo; units =
 * 16-column blocks over the batch.  1D n = 65536 has no list (its own
 * two-pass kernel): *n_passes = 0.  capacity < passes: TCFFT_INVALID_VALUE
 * with *n_passes set.  No counterpart in the reference. */
int tcfft_query_split(int dims, int n_or_nx, int ny, int n_batch, int *desc, int capacity, int *n_passes);

/* Real-input transforms (no counterpart in the reference, whose callers feed
 * real data as complex with zero imaginary parts, e.g. the generator of
 * doit.h:7-19 fills both).  A plan for n_batch real signals of n = 2^9 ..
 * 2^27 points (else TCFFT_INVALID_SIZE); it is also the complex plan of n/2
 * points for every other entry point.  Up to n = 2^16 each direction is one
 * HBM round trip of the half-size data (the even/odd split fused into the
 * transform).  Above, the n/2-point transform runs as the plan's passes in
 * natural order through a workspace of n * n_batch * 2 bytes that the plan
 * allocates at creation, and the split is a pass of its own; concurrent
 * calls on one such plan race on that workspace.
 *   tcfft_exec_r2c: in = n_batch * n FP16 reals (16-byte aligned), out =
 *     bins 0..n/2 of each signal's DFT (sign e^{-2 pi i nk/N}), n/2 + 1 FP16
 *     complex per signal, signal b at out + 4*(n/2+1)*b bytes (4-byte aligned);
 *   tcfft_exec_c2r: in = n/2 + 1 bins per signal as above, taken as the
 *     lower half of a Hermitian spectrum (imaginary parts of bins 0 and n/2
 *     ignored), out = y[t] = sum_{k<n} X[k] e^{+2 pi i kt/n} (unnormalised,
 *     n times numpy's irfft), n FP16 reals per signal.
 * Out of place only (the buffers must not overlap: TCFFT_INVALID_POINTER);
 * no allocation in the exec calls, graph-capturable.  A plan not made by
 * tcfft_plan_create_r2c_1d gives TCFFT_INVALID_PLAN. */
int tcfft_plan_create_r2c_1d(tcfft_plan *plan, int n, int n_batch);
int tcfft_exec_r2c(tcfft_plan plan, const void *in, void *out, void *stream);
int tcfft_exec_c2r(tcfft_plan plan, const void *in, void *out, void *stream);

/* Natural-order input, spectrum output in the plan's digit-reversed order
 * (position p holds X[rev[p]], rev = tcfft_gen_rev of the plan's radices):
 * the order tcfft_exec consumes, so a transform, a pointwise product in that
 * order and an inverse tcfft_exec need no permutation.  1D plans with
 * n = 4096, 8192 or 16384; in place or out of place.  No reference
 * counterpart (the reference takes digit-reversed input only); it replaces
 * the caller's host-side gen_rev staging (tcfft_doit_half.cpp:40-46) for
 * spectral callers. */
int tcfft_exec_dif(tcfft_plan plan, const void *in, void *out, int direction, void *stream);

/* Fused FFT convolution along the transform axis, one HBM round trip:
 *   y[b][c] = F^H (H[c] * F x[b][c]) / n      (conj_filter = 0)
 *   y[b][c] = F^H (conj(H[c]) * F x[b][c]) / n (conj_filter = 1, the data
 *                                                gradient of the former)
 * x, y: [batch/channels][channels][n] fp16 complex, natural order (x == y
 * allowed); h_rev: [channels][n] fp16 complex spectrum in the digit-reversed
 * order above; spec_out (nullable): receives F x, digit-reversed.  The plan
 * is a 1D plan of n = 4096, 8192 or 16384 whose batch counts all signals.
 * No reference counterpart (an FFT-conv layer built on the transform). */
int tcfft_conv_exec(tcfft_plan plan, int channels, const void *x, const void *h_rev, void *y, int conj_filter,
                    void *spec_out, void *stream);

/* Filter gradient of tcfft_conv_exec: gh_rev[c][p] = (1/n) sum_b
 * conj(X[b][c][p]) * G[b][c][p] from digit-reversed fp16 spectra of the
 * input and of the output gradient; gh_rev is FP32 complex [channels][n],
 * digit-reversed.  No reference counterpart. */
int tcfft_conv_filter_grad(tcfft_plan plan, int channels, const void *x_spec, const void *g_spec, void *gh_rev,
                           void *stream);

/* Plan options. TCFFT_OPT_SPLIT_CHUNK: signals per pass pair of
 * tcfft_exec_split (0 = whole batch, the default), which bounds the
 * intermediate's footprint (smaller chunks keep it in the last-level cache
 * but leave compute units idle; at 2^16 x 256 the whole batch is fastest).
 * TCFFT_OPT_BIG_DELAY (-1..64, default -1 = per-kernel default; the
 * natural-order 65536-point kernel then runs 5 groups with delay 3) and
 * TCFFT_OPT_BIG_GROUPS (1..8, default 3): start schedule of the first pass
 * (contiguous lines of up to 65536 points); the workgroups resident at
 * launch form GROUPS groups and group k starts k * DELAY quanta (1024 GPU
 * cycles each) late, so that one group's loads overlap another's stores
 * instead of all workgroups loading, computing and storing in lockstep.
 * TCFFT_OPT_COLS_DELAY (0..64, default 1): the same start delay (3 groups)
 * for a strided pass of radix 512..2048.
 * TCFFT_OPT_SPLIT_DELAY (0..64, default 2): the same start delay (3 groups)
 * for the two passes of tcfft_exec_split.
 * TCFFT_OPT_ROWS_WAVES (1, 2 or 4; 0, the default, = 4 for 4096-point lines
 * over at least 32 waves per compute unit, else 1): waves per workgroup of a
 * contiguous-row first pass (lines of 256..4096 points, one wave per 16
 * tiles of 256 points), i.e. how the waves are grouped onto compute units.
 * TCFFT_OPT_FIRST_CACHED (-1, the default: cached while the signals of one
 * batch chunk take at most 256 MiB; 0 = never; 1 = always): whether the first pass
 * of a multi-pass plan stores its output through the caches, so that the
 * next pass finds much of it in the 256 MiB last-level cache (MALL), instead
 * of with non-temporal stores.
 * TCFFT_OPT_MALL_CHUNK (MiB, 16..65536; 0 = off; default 256): a multi-pass
 * transform whose buffer exceeds this runs all its passes on one chunk of
 * whole signals (images) of at most this size before the next, so that the
 * later passes read the earlier ones' output from the MALL.  Plans with a
 * radix-512..2048 column pass (2D nx = 512..2048, 1D 2^25..2^27) and batches
 * whose single signal exceeds the size are not chunked.
 * TCFFT_OPT_ROWS_TILES (4, 8 or 16; 0, the default, = by the grid): tiles of
 * 256 points per wave of a contiguous-row first pass without a folded
 * twiddle (at least the line's tiles; for 4096-point lines 8 and 4 mean two
 * and four waves per line).  The default is the largest that still gives
 * every SIMD two waves (16 on all but small batches); for 4096-point lines
 * 4 up to one line per CU, 8 up to four, 16 above, and on natural-order
 * input 4 up to one line per CU, 16 from 4 to 8 lines per CU, 8 otherwise.
 * None of these options changes results. */
enum tcfft_option {
    TCFFT_OPT_SPLIT_CHUNK = 1,
    TCFFT_OPT_BIG_DELAY = 2,
    TCFFT_OPT_BIG_GROUPS = 3,
    TCFFT_OPT_COLS_DELAY = 4,
    TCFFT_OPT_SPLIT_DELAY = 5,
    TCFFT_OPT_ROWS_WAVES = 6,
    TCFFT_OPT_FIRST_CACHED = 7,
    TCFFT_OPT_MALL_CHUNK = 8,
    TCFFT_OPT_ROWS_TILES = 9
};
int tcfft_plan_set_option(tcfft_plan plan, int option, int value);

/* Times the plan's forward transform under each start schedule candidate
 * of its first pass (no stagger, or 3, 5 or 7 groups with delays of 1..4
 * quanta) on `data`, keeps the fastest in the plan and reports it; then,
 * with that schedule, the workgroup size of a contiguous-row first pass
 * (TCFFT_OPT_ROWS_WAVES), its tiles per wave (TCFFT_OPT_ROWS_TILES; lines
 * up to 2048 points, 4096-point lines keep the choice by the grid) and the
 * column pass delay (0..4) when the plan has a strided pass of radix
 * 512..2048.  Every timed run starts from the caller's data (a copy that
 * the call allocates: transforms repeated in place would time inf/NaN
 * values, which run faster than finite data) and follows a write of a
 * 256 MiB buffer that the call allocates and frees, so that the data is
 * timed cold, as a caller's previous work leaves it; the four fastest
 * candidates are timed again in three alternating rounds before the choice.
 * On return `data` holds what it held on entry (if the copy cannot be
 * allocated, transforms of itself, and the timed runs see those).  No
 * reference counterpart (the reference has one fixed launch per size). */
int tcfft_plan_autotune(tcfft_plan plan, void *data, void *stream, int *best_groups, int *best_delay,
                        float *best_ms);

/* Uneven batches (no counterpart in the reference): n_signals transforms
 * of sizes[i] (powers of two, 2^8 .. 2^16) packed back to back in one
 * buffer (signal i at element offset sizes[0] + ... + sizes[i-1]), each in
 * the drop-in order (digit-reversed by the radix list a 1D plan of that size
 * reports), transformed in place by one persistent launch.  Workgroups walk
 * equal-sized tasks (one size class, <= 65536 elements each) round-robin. */
typedef struct tcfft_multi_s *tcfft_multi;
int tcfft_multi_create(tcfft_multi *plan, int n_signals, const int *sizes);
int tcfft_multi_destroy(tcfft_multi plan);
int tcfft_multi_info(tcfft_multi plan, int *n_tasks, long long *total_elements);
int tcfft_multi_exec(tcfft_multi plan, void *data, int direction, void *stream);
/* The same transforms on natural-order signals (no counterpart in the
 * reference, whose callers permute on the host, tcfft_doit_half.cpp:40-46):
 * signal i of `in` at its packed offset, in natural order, to the same place
 * in `out`, natural order.  in == out (in place) is allowed; in and out must
 * not otherwise overlap.  Bit-identical to tcfft_multi_exec on each signal
 * permuted by its own plan's gen_rev.  One persistent launch (the plan's
 * schedule); the digit reversal runs in the tasks' loads through LDS.
 * Allocates nothing; graph-capturable. */
int tcfft_multi_exec_natural(tcfft_multi plan, const void *in, void *out, int direction, void *stream);
/* Schedule: grid = persistent workgroups (0 = one per task), order = task
 * order (0 largest sizes first, 1 interleaved, 2 smallest first). */
int tcfft_multi_set_schedule(tcfft_multi plan, int grid, int order);
/* Start delay of the resident workgroups (0..64; 0 = none): they form 3
 * groups by their index within an XCD and group k starts k * delay quanta of
 * 1024 GPU cycles late, so that loads and stores of different groups overlap.
 * Never changes results. */
int tcfft_multi_set_start_delay(tcfft_multi plan, int delay);
int tcfft_multi_get_schedule(tcfft_multi plan, int *grid, int *order, int *delay);
/* Times every (grid, order, start delay 0..2) candidate on `data`, times the
 * four fastest again in three alternating rounds, and keeps the one with the
 * lowest time.  As in tcfft_plan_autotune, each timed run starts from a copy
 * of the caller's data and follows a write of a 256 MiB buffer (each
 * skipped if its allocation fails), and `data` is restored on return. */
int tcfft_multi_autotune(tcfft_multi plan, void *data, void *stream, int *best_grid, int *best_order,
                         float *best_ms);

/* Runs passes [first, first + count) of the plan only, each over the whole
 * batch (the buffer must hold what the earlier passes produced).  For
 * profiling individual pass kernels.  The result of all passes equals
 * tcfft_exec's bit for bit; tcfft_exec may run them per batch chunk
 * (TCFFT_OPT_MALL_CHUNK). */
int tcfft_plan_num_passes(tcfft_plan plan, int *n_passes);
int tcfft_exec_passes(tcfft_plan plan, void *data, int direction, int first, int count, void *stream);

/* Host helper with the semantics of the caller-side gen_rev
 * (tcfft_doit_half.cpp:6-29): rev[i] = source index of position i. */
int tcfft_gen_rev(int n, const int *radices, int n_radices, int *rev);

#ifdef __cplusplus
}
#endif
#endif
