// fft_kernels.hpp -- launch interface between the pass driver and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcfft {

// Twiddle that a pass applies to its FP32 results for the pass that follows
// (see epi_twiddle in fft_kernels.hip).  hi/lo are FP64-generated FP32 tables
// of exp(-2*pi*i*k/M) split as W_M^m = hi[m >> lo_bits] * lo[m & (2^lo_bits-1)].
struct EpiParams {
    const float2* hi;
    const float2* lo;
    uint32_t lo_bits;
    uint32_t m_mask;   // M - 1, M = s * R of the next pass
    uint32_t shift;    // axis position = element index >> shift
    uint32_t s_log2;   // next pass step, in axis positions
    uint32_t r_log2;   // next pass radix
    uint32_t ra_log2;  // first radix inside the next pass (DIT order)
    uint32_t rb_log2;  // second radix inside the next pass (the rest is the third)
};

// Start schedule of the first-pass kernels (k_rows, k_big).  Workgroups of
// the first resident wave (b < resident) are put in `groups` groups by their
// index within an XCD (b >> 3, blocks go round-robin over the 8 XCDs); group
// k sleeps k * delay quanta of 1024 cycles before its first load, so the
// groups' load, compute and store phases overlap instead of running in
// lockstep.  delay = 0: none.
struct StartSched {
    uint32_t delay;
    uint32_t groups;
    uint32_t resident;
    // Not a schedule but a launch policy carried with it: 1 = the pass's
    // output stores go through the caches (k_big / k_big_nat, and k_rows /
    // k_rows_nat for 4096-point lines, whose stores are otherwise
    // non-temporal; a separate kernel instance).  Set for a first pass that
    // a further pass reads back while the buffer fits in the MALL.
    uint32_t cached_out;
};

// First pass: contiguous lines of 256*T elements (T = 1, 2, 4, 8, 16), one
// wave per tiles_per_wave tiles (16; without an epilogue also 8, or 4 when
// T <= 4; for T = 16, 8 or 4 is two or four waves per line), wg_waves (1,
// 2 or 4) waves per workgroup (two-wave lines: 4 = two lines per workgroup,
// else one).
hipError_t launch_rows(int T, uint32_t* buf, uint32_t n_tiles, int sign, const float2* w4096,
                       const EpiParams* epi, StartSched sched, uint32_t wg_waves, uint32_t tiles_per_wave,
                       hipStream_t st);

// Single-pass transform of N = 4096*W (W = 2, 4, 8, 16) per workgroup of W
// waves.  wn_hi/wn_lo: two-level table of W_N (lo_bits split).
hipError_t launch_big(int W, uint32_t* buf, uint32_t n_signals, int sign, const float2* w4096, const float2* wn_hi,
                      const float2* wn_lo, uint32_t lo_bits, const EpiParams* epi, StartSched sched, hipStream_t st);

// Last strided pass of its axis with radix R = 256*W (W = 2, 4, 8) at element
// step 2^s_log2 over `outer` [R][s] blocks, in one HBM round trip.
hipError_t launch_cols_big(int W, uint32_t* buf, uint32_t s_log2, uint32_t outer, int sign, const float2* w4096,
                           StartSched sched, hipStream_t st);

// Split-precision N = 65536 transform of FP32 complex data, as two passes per
// chunk of `chunk` signals (0 = all at once): in -> mid (transposed), then
// mid -> out (mid may be out; in must differ from mid).  `sched` = the
// passes' start schedule (delay and groups; the resident count is filled in).
hipError_t launch_split_65536(const float2* in, float2* mid, float2* out, uint32_t n_signals, uint32_t chunk,
                              int sign, const float2* w4096, const float2* wn_hi, const float2* wn_lo,
                              StartSched sched, hipStream_t st);

// One split-precision column pass (split.hip): each signal is R rows x Q
// columns (element (r, q) at r*Q + q); the R-point DFT down every column,
// output k of column q = q_hi*Qlo + q_lo times W_M^{q_hi k} (M = R*Q/Qlo;
// m_mask = 0: no twiddle), stored at q_hi*R*Qlo + k*Qlo + q_lo.  Qlo = 1 is
// the row store (row = true); otherwise Qlo >= 16.  R = 16..256 on MFMA,
// R = 2, 4, 8 (no twiddle, Qlo = Q) on VALU.  src != dst unless Qlo = Q.
struct SplitPassArgs {
    uint32_t q_log2;    // Q = 2^q_log2 >= 16
    uint32_t qlo_log2;  // Qlo = 2^qlo_log2
    uint32_t units;     // 16-column units: signals * Q / 16
    uint32_t m_mask;    // M - 1, 0 = no twiddle
    uint32_t lo_bits;   // W_M^m = hi[m >> lo_bits] * lo[m & (2^lo_bits - 1)]
    const float2* hi;
    const float2* lo;
};
hipError_t launch_split_pass(int R, bool row, const float2* src, float2* dst, const SplitPassArgs& a, int sign,
                             const float2* w4096, StartSched sched, hipStream_t st);

// Uneven batch: one task = signals [first, first + count) of one size class.
struct MultiTask {
    uint32_t size_log2;
    uint32_t first;
    uint32_t count;
    uint32_t pad;
};
struct MultiTables {  // two-level W_N tables for N = 2^13 .. 2^16
    const float2* hi[4];
    const float2* lo[4];
};
hipError_t launch_multi(uint32_t* buf, const MultiTask* tasks, uint32_t n_tasks, const uint32_t* offsets,
                        uint32_t grid, int sign, const float2* w4096, const MultiTables& tb, StartSched sched,
                        hipStream_t st);
// The same tasks on natural-order input: in -> out (in == out allowed).
hipError_t launch_multi_nat(const uint32_t* in, uint32_t* out, const MultiTask* tasks, uint32_t n_tasks,
                            const uint32_t* offsets, uint32_t grid, int sign, const float2* w4096,
                            const MultiTables& tb, StartSched sched, hipStream_t st);

// Natural-order transform of N = 4096*W (W = 1, 2, 4) per signal whose output
// is the spectrum in the digit-reversed order the DIT kernels consume.
hipError_t launch_dif(int W, const uint32_t* in, uint32_t* out, uint32_t n_signals, int sign, const float2* w4096,
                      const float2* wn_hi, const float2* wn_lo, uint32_t lo_bits, hipStream_t st);
// Fused FFT convolution y = F^H (H * F x) * scale (N = 4096*W, W = 1, 2, 4);
// h_rev digit-reversed fp16 spectrum per channel, conj -> conj(H); spec
// (nullable) receives F x, digit-reversed.
hipError_t launch_conv(int W, bool conj, const uint32_t* x, const uint32_t* h_rev, uint32_t* y, uint32_t* spec,
                       uint32_t n_signals, uint32_t channels, float scale, const float2* w4096, const float2* wn_hi,
                       const float2* wn_lo, uint32_t lo_bits, hipStream_t st);
// gh_rev[c][p] = scale * sum_b conj(X[b][c][p]) G[b][c][p] (FP32 out).
hipError_t launch_conv_wgrad(const uint32_t* xs, const uint32_t* gs, float2* gh, uint32_t n, uint32_t channels,
                             uint32_t batch, float scale, hipStream_t st);

// Strided pass of radix R (2..256) at element step 2^s_log2 over `outer` blocks.
hipError_t launch_cols(int R, uint32_t* buf, uint32_t s_log2, uint32_t outer, int sign, const float2* w4096,
                       const EpiParams* epi, hipStream_t st);

// Digit-reversal gather (natural-order input).  rev digits are described by
// the radix list (up to 12 digits per axis).
struct RevSpec {
    int n_digits;
    int radix[12];
};
// First pass of k_rows<T> on natural-order input, in -> out (in place allowed
// when nx_log2 = 0).  2D (nx_log2 > 0, out of place): input line img*nx + x is
// written to line img*nx + xpos[x] (xpos: device table of nx words, the digit
// reversal over the x radices in reverse order).  tiles_per_wave: 16, or 8
// or 4 for T = 16 (two or four waves per line).
hipError_t launch_rows_nat(int T, const uint32_t* in, uint32_t* out, uint32_t n_tiles, int sign, const float2* w4096,
                           uint32_t nx_log2, const uint32_t* xpos, StartSched sched, uint32_t tiles_per_wave,
                           hipStream_t st);

// Real-input transforms of N = 512*T points (T = 1, 2, 4, 8, 16), n_tiles =
// batch * T.  sign -1: real in (batch * N/2 words, 16-byte aligned) -> bins
// 0..N/2 out (batch * (N/2 + 1) words); sign +1: the inverse.  wn = W_N^0..64.
hipError_t launch_r2c(int T, int sign, const uint32_t* in, uint32_t* out, uint32_t n_tiles, const float2* w4096,
                      const float2* wn, hipStream_t st);
// The even/odd split alone for N = 2L > 65536 (L = 2^l_log2): sign -1 maps
// the natural L-point spectrum Z (L words per signal) to bins 0..L (L+1 words
// per signal), sign +1 the bins to Z' (L words per signal).  hi/lo/lo_bits:
// the two-level table of W_N.
hipError_t launch_real_pairs(int sign, const uint32_t* in, uint32_t* out, uint32_t l_log2, uint32_t n_signals,
                             const float2* hi, const float2* lo, uint32_t lo_bits, hipStream_t st);
// Same for N = 8192*W (W = 2, 4, 8), one workgroup per signal; wn_hi/wn_lo/
// lo_bits = the two-level W_{4096W} table of the half-size plan's first pass,
// wn = W_N^0..64W.
hipError_t launch_r2c_big(int W, int sign, const uint32_t* in, uint32_t* out, uint32_t n_signals,
                          const float2* w4096, const float2* wn_hi, const float2* wn_lo, uint32_t lo_bits,
                          const float2* wn, StartSched sched, hipStream_t st);

// Tile of the LDS-tiled 1D digit reversal: a low bits of the position and
// the b position bits in_bit[0..b-1] that form the low bits of the input
// index (disjoint, a + b <= 12); free_mask = the remaining position bits.
struct GatherTile {
    uint32_t a, b;
    uint32_t free_mask;
    uint32_t in_bit[6];      // position bit giving bit i of the input index, i < b
    uint32_t low_to_rev[6];  // input-index bit of position bit j, j < a
};
hipError_t launch_gather_tiled_1d(const uint32_t* in, uint32_t* out, uint32_t n_log2, uint32_t total,
                                  const RevSpec& spec, const GatherTile& gt, hipStream_t st);

// Single-pass transform of N = 4096*W (W = 2, 4, 8, 16) on natural-order
// input, in -> out (in place allowed when nx_log2 = 0); 2D as launch_rows_nat.
hipError_t launch_big_nat(int W, const uint32_t* in, uint32_t* out, uint32_t n_signals, int sign, const float2* w4096,
                          const float2* wn_hi, const float2* wn_lo, uint32_t lo_bits, uint32_t nx_log2,
                          const RevSpec& x_pos, StartSched sched, hipStream_t st);

}  // namespace tcfft
