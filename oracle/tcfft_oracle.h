/*
 * tcfft_oracle.h -- CPU restatement of tcFFT's half-precision FFT path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (tcfft_amd/).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it and has no CPU
 * fallback onto it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Data layout everywhere: interleaved complex doubles, signal b of length n
 * at data + 2*n*b (re at even, im at odd offsets), as in doit.h:7-19.
 *
 * Pinning (see tests/test_oracle.py):
 *   oracle_gen_rev         closed-form mixed-radix digit reversal, and
 *                          "gen_rev + oracle_dit == brute-force DFT"
 *   oracle_generate_data   an independent pure-Python model of glibc
 *                          random_r (TYPE_3) plus the srand(1) known answer
 *   oracle_get_error       hand-computed cases (clamp at 1, x/0, exact)
 *   oracle_ref_plan_1d/2d  every listed radix list reproduces the DFT through
 *                          gen_rev + oracle_dit (the reference's own contract)
 *   oracle_dit / _2d       brute-force long-double DFT, numpy pocketfft golden
 *                          vectors, Parseval, impulse, tone, linearity,
 *                          double-transform identity, 2D separability
 *   oracle_fft / _2d       same as oracle_dit (natural order wrapper)
 *   oracle_dft_ld          textbook closed forms (impulse, constant, tone) and
 *                          the numpy golden vectors
 *   oracle_rel_l2          hand-computed cases
 *   oracle_rfft / _irfft   numpy rfft / irfft, Hermitian symmetry of the full
 *                          transform, real-cosine and Nyquist closed forms,
 *                          irfft(rfft(x)) = n x
 */
#ifndef TCFFT_ORACLE_H
#define TCFFT_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Caller-side digit reversal: rev[i] is the source index of position i.
 * tcfft_doit_half.cpp:6-29 (applied per axis in tcfft_doit_half_2d.cpp:40-41). */
void oracle_gen_rev(int n, int *rev, const int *radices, int n_radices);

/* Reference input generator: glibc srand(seed) / rand(), values
 * 0.0001f*rand()/RAND_MAX for re then im.  doit.h:7-19. */
void oracle_generate_data(double *data, int n, int n_batch, int seed);

/* Reference accuracy statistic: mean over all re/im components of
 * min(1, |(t-s)/s|).  accuracy.cpp:25-41. */
double oracle_get_error(const double *tested, const double *standard, int n, int n_batch);

/* Reference plan table.  Fills the radix list and pass ("merging") list that
 * tcfftCreate builds for n (tcfft_half.cu:1235-1330).  Returns 0 for the ten
 * sizes the table lists, -1 for any other n (the reference then runs a wrong
 * transform, see tcfft_half.cu:1328-1329). */
int oracle_ref_plan_1d(int n, int *radices9, int *n_radices, int *mergings3, int *n_mergings);

/* 2D reference plan table, tcfft_half_2d.cu:1151-1206.  Returns 0 when both
 * axes are listed and the column pass has a kernel (nx in {256,512}),
 * -1 otherwise (nx=1024 selects the unassigned layer_1[2]). */
int oracle_ref_plan_2d(int nx, int ny, int *radices_x3, int *n_radices_x,
                       int *radices_y3, int *n_radices_y, int *mergings2);

/* In-place mixed-radix decimation-in-time FFT on digit-reversed input with
 * natural-order output: the arithmetic of tcfftExec's pass sequence
 * (tcfft_half.cu:1177-1233) and its pass kernels (layer_*_0 / layer_*_1,
 * tcfft_half.cu:69-161, 266-390, 585-699, 774-883, 1010-1095).  Stage i
 * combines radix radices[i] at step r_0*...*r_{i-1}; twiddles W_{s*r}^{c*j}
 * are exact FP64 values.  sign = -1 is FFTW_FORWARD (W_N_K, tcfft_half.cu:40-44);
 * sign = +1 is the unnormalised conjugate (adjoint) transform. */
void oracle_dit(double *data, int n, int n_batch, const int *radices, int n_radices, int sign);

/* 2D: rows (length ny, contiguous) then columns (length nx, stride ny), both
 * digit-reversed on input.  tcfft_half_2d.cu:1128-1149 with the per-axis
 * reversal of tcfft_doit_half_2d.cpp:40-50.  The column pass has no
 * cross-dimension twiddle (tcfft_half_2d.cu:305-315). */
void oracle_dit_2d(double *data, int nx, int ny, int n_batch,
                   const int *radices_x, int n_radices_x,
                   const int *radices_y, int n_radices_y, int sign);

/* Natural-order FP64 FFT (the role FFTW double plays in accuracy.cpp:11-23 and
 * accuracy_2d.cpp:11-23): gen_rev permutation + oracle_dit with a
 * radix-16-first list.  n must be a power of two.  in and out may alias. */
void oracle_fft(const double *in, double *out, int n, int n_batch, int sign);
void oracle_fft_2d(const double *in, double *out, int nx, int ny, int n_batch, int sign);

/* Real-input transform, bins 0..n/2 of each signal's DFT (n/2 + 1 complex
 * per signal; in: n_batch * n doubles).  No counterpart in the reference,
 * whose callers feed real data as complex with zero imaginary parts (the
 * FFT-conv workload; doit.h:7-19 fills both components): this is exactly
 * that, oracle_fft of the real-as-complex input, truncated to n/2 + 1 bins. */
void oracle_rfft(const double *in, double *out, int n, int n_batch);

/* Inverse of the above, unnormalised: y[t] = sum_{k<n} X[k] e^{+2 pi i kt/n}
 * over the Hermitian extension X[n-k] = conj X[k] of bins 0..n/2, with the
 * imaginary parts of bins 0 and n/2 dropped (n * numpy's irfft). */
void oracle_irfft(const double *in, double *out, int n, int n_batch);

/* O(n^2) DFT in long double with exactly reduced angles (pin for the above). */
void oracle_dft_ld(const double *in, double *out, int n, int sign);

/* ||a - ref||_2 / ||ref||_2 over count doubles. */
double oracle_rel_l2(const double *a, const double *ref, size_t count);

/* Radix-16-first radix list for a power-of-two n (used by oracle_fft). */
int oracle_default_radices(int n, int *radices, int capacity);

#ifdef __cplusplus
}
#endif
#endif
