// fft_kernels.hip -- gfx950 pass kernels for the FP16 FFT.
//
// Every pass is one step of a mixed-radix decimation-in-time FFT on
// digit-reversed input (the contract of tcfftExec, tcfft_half.cu:1177-1233):
// a pass of radix R at step s combines, for every column c < s of every block
// of s*R elements, the R rows j*s + c.  The twiddle W_{s*R}^{c*rev(j)} that
// precedes a pass is applied by the previous pass in FP32 before it rounds to
// FP16 (the "epilogue twiddle"), so a pass reads FP16 straight into MFMA.
//
//   k_rows<T>       first pass: contiguous lines of L = 256*T (T = 1..16),
//                   one wave owns 16 tiles of 256 elements in registers.
//                   Radix 16 x 16 on MFMA (A-form twice, no LDS), then a
//                   radix-T stage across tiles (VALU for T = 2,4,8; MFMA after
//                   a permlane register transpose for T = 16).
//   k_cols<T>       strided pass of radix 16*T (T = 1..16) over 16 columns
//                   per wave (B-form MFMA; T = 16 uses the permlane transpose).
//   k_cols_small<R> strided pass of radix 2, 4 or 8 in FP32 VALU, four
//                   adjacent columns per lane.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace tcfft {

// Epilogue twiddle for the next pass: element at position P (within the
// transform axis, P = global_pos >> shift) gets W_{s*R}^{c*rev(j)} with
// c = P mod s, j = (P / s) mod R and rev the two-digit DIT reversal of the
// next pass (radix ra first, then R/ra).  W_M^m = hi[m >> lo_bits]*lo[m & lo_mask].
__device__ __forceinline__ cf epi_twiddle(const EpiParams& e, u32 gpos, float sgn) {
    const u32 P = gpos >> e.shift;
    const u32 c = P & ((1u << e.s_log2) - 1u);
    const u32 j = (P >> e.s_log2) & ((1u << e.r_log2) - 1u);
    const u32 j0 = j & ((1u << e.ra_log2) - 1u);
    const u32 j1 = j >> e.ra_log2;
    const u32 rj = j1 | (j0 << (e.r_log2 - e.ra_log2));
    const u32 m = (c * rj) & e.m_mask;
    const cf a = twiddle(e.hi, m >> e.lo_bits, sgn);
    const cf b = twiddle(e.lo, m & ((1u << e.lo_bits) - 1u), sgn);
    return cmul(a, b);
}

// ----------------------------------------------------------------------------
// k_rows<T>: lines of L = 256*T contiguous elements, digit-reversed for the
// radix list [16, 16] (+ [T] when T > 1).  Position in a line:
// p = d0 + 16*d1 + 256*d2.  Wave layout at load: tile t = 256-element block,
// lane (c16 = d1, g) holds d0 = 4g..4g+3 in its four words (one 16-byte load).
template <int T, bool EPI>
__global__ __launch_bounds__(256) void k_rows(u32* __restrict__ buf, u32 n_tiles, int sign,
                                              const float2* __restrict__ w4096, EpiParams epi) {
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g = lane >> 4;
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const u32 tile0 = wave * 16u;
    if (tile0 >= n_tiles) return;
    const u32 ntl = min(16u, n_tiles - tile0);
    const float sg = sign < 0 ? 1.f : -1.f;

    const Dft16Op op = make_dft16(lane, sign);
    cf tw1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) tw1[i] = twiddle(w4096, ((c16 * (4 * g + i)) & 255) * 16, sg);

    u32 d[16][4];
    u32* const base = buf + (size_t)tile0 * 256 + 16 * c16 + 4 * g;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        if ((u32)t < ntl) load4(d[t], base + t * 256);
        else d[t][0] = d[t][1] = d[t][2] = d[t][3] = 0u;
    }

    // Stage 0: DFT over d0 (A-form) -> lane%16 = k0, rows (g,i) = d1.
    // Twiddle W_256^{k0*d1} is a per-lane constant.
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(cmul(cf{re[i], im[i]}, tw1[i]));
    }

    if constexpr (T == 1) {
        // Stage 1: DFT over d1 -> lane%16 = k1, rows (g,i) = k0; natural
        // position c = k0 + 16*k1 = 4g + i + 16*c16: four adjacent words.
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_a(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {re[i], im[i]};
                if constexpr (EPI) {
                    const u32 gpos = (tile0 + t) * 256u + 16u * c16 + 4u * g + i;
                    v = cmul(v, epi_twiddle(epi, gpos, sg));
                }
                d[t][i] = pack(v);
            }
            if ((u32)t < ntl) store4(base + t * 256, d[t]);
        }
    } else if constexpr (T < 16) {
        // Stage 1 on each tile of a line, then radix-T across the line's tiles:
        // X[c + 256*k2] = sum_d2 W_{256T}^{c*d2} Y_d2[c] W_T^{d2*k2}.
        constexpr int STRIDE = 16 / T;  // W_{256T}^m = W_4096^{m*STRIDE}
#pragma unroll
        for (int gi = 0; gi < 16 / T; ++gi) {
            cf acc[T][4];
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                f32x4 re, im;
                dft16_a(d[gi * T + tt], op, re, im);
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[tt][i] = cf{re[i], im[i]};
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u32 c = 16u * c16 + 4u * g + i;
                cf v[T];
                v[0] = acc[0][i];
#pragma unroll
                for (int tt = 1; tt < T; ++tt)
                    v[tt] = cmul(acc[tt][i], twiddle(w4096, ((c * tt) * STRIDE) & 4095u, sg));
                small_dft<T>(v, sign);
#pragma unroll
                for (int tt = 0; tt < T; ++tt) acc[tt][i] = v[tt];
            }
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                const int t = gi * T + tt;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    cf v = acc[tt][i];
                    if constexpr (EPI) {
                        const u32 gpos = (tile0 + t) * 256u + 16u * c16 + 4u * g + i;
                        v = cmul(v, epi_twiddle(epi, gpos, sg));
                    }
                    d[t][i] = pack(v);
                }
                if ((u32)t < ntl) store4(base + t * 256, d[t]);
            }
        }
    } else {
        // T == 16: one 4096-point line per wave.
        // Stage 1 -> lane%16 = k1, rows (g,i) = k0, tile = d2; then the
        // stage-2 twiddle W_4096^{(k0+16k1)*d2} in FP32 before rounding.
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_a(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u32 c = 16u * c16 + 4u * g + i;
                cf v = {re[i], im[i]};
                if (t) v = cmul(v, twiddle(w4096, (c * t) & 4095u, sg));
                d[t][i] = pack(v);
            }
        }
        // Register transpose: (g,i) = k0, tile = d2  ->  (g,i) = d2, tile = k0.
        u32 e[16][4];
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) e[(t & 12) | i][t & 3] = d[t][i];
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 4))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l4(e[t][i], e[t | 4][i]);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 8))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l5(e[t][i], e[t | 8][i]);
        // Stage 2 (B-form): columns k1 = lane%16, points d2 = (g,i) -> rows (g,i) = k2.
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(e[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {re[i], im[i]};
                if constexpr (EPI) {
                    const u32 k = (u32)t + 16u * c16 + 256u * (4u * g + i);
                    v = cmul(v, epi_twiddle(epi, tile0 * 256u + k, sg));
                }
                e[t][i] = pack(v);
            }
        }
        // Register transpose for the store: (g,i) = k2, tile = k0 -> (g,i) = k0, tile = k2.
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) d[(t & 12) | i][t & 3] = e[t][i];
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 4))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l4(d[t][i], d[t | 4][i]);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 8))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l5(d[t][i], d[t | 8][i]);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if ((u32)t < ntl) store4(base + t * 256, d[t]);
    }
}

// ----------------------------------------------------------------------------
// k_cols<T>: strided pass of radix R = 16*T at element step s over a buffer
// viewed as [outer][R][s].  A wave owns 16 adjacent columns of one block.
// Rows j = j0 + 16*j1 are digit-reversed for [16, T]: stage 0 is a DFT over
// j0 (adjacent rows), stage 1 a radix-T over j1 with W_{16T}^{k0*j1}.
// Lane (c16, g) word i of tile j1 holds row j0 = 4g+i, column c16.
template <int T, bool EPI>
__global__ __launch_bounds__(256) void k_cols(u32* __restrict__ buf, u32 s_log2, u32 n_waves, int sign,
                                              const float2* __restrict__ w4096, EpiParams epi) {
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g = lane >> 4;
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= n_waves) return;
    const u32 s = 1u << s_log2;
    const u32 col_groups = s >> 4;
    const u32 blk = wave / col_groups;
    const u32 col0 = (wave - blk * col_groups) * 16u;
    const float sg = sign < 0 ? 1.f : -1.f;
    const Dft16Op op = make_dft16(lane, sign);
    constexpr u32 R = 16 * T;
    const size_t blk_base = (size_t)blk * R * s + col0 + c16;

    u32 d[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = buf[blk_base + (size_t)(16 * t + 4 * g + i) * s];

    auto out_pos = [&](u32 row) -> u32 { return (u32)(blk_base + (size_t)row * s); };

    if constexpr (T == 1) {
        f32x4 re, im;
        dft16_b(d[0], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf v = {re[i], im[i]};
            const u32 row = 4u * g + i;
            if constexpr (EPI) v = cmul(v, epi_twiddle(epi, out_pos(row), sg));
            buf[blk_base + (size_t)row * s] = pack(v);
        }
    } else if constexpr (T < 16) {
        constexpr int STRIDE = 256 / T;  // W_{16T}^m = W_4096^{m*STRIDE}
        cf acc[T][4];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[t][i] = cf{re[i], im[i]};
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const u32 k0 = 4u * g + i;
            cf v[T];
            v[0] = acc[0][i];
#pragma unroll
            for (int t = 1; t < T; ++t) v[t] = cmul(acc[t][i], twiddle(w4096, ((k0 * t) * STRIDE) & 4095u, sg));
            small_dft<T>(v, sign);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const u32 row = k0 + 16u * t;
                cf o = v[t];
                if constexpr (EPI) o = cmul(o, epi_twiddle(epi, out_pos(row), sg));
                buf[blk_base + (size_t)row * s] = pack(o);
            }
        }
    } else {
        // T == 16: stage 0 -> (g,i) = k0, tile = j1; twiddle W_256^{k0*j1};
        // transpose to (g,i) = j1, tile = k0; stage 1 -> (g,i) = k1.
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {re[i], im[i]};
                const u32 k0 = 4u * g + i;
                if (t) v = cmul(v, twiddle(w4096, ((k0 * t) & 255u) * 16u, sg));
                d[t][i] = pack(v);
            }
        }
        u32 e[16][4];
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) e[(t & 12) | i][t & 3] = d[t][i];
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 4))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l4(e[t][i], e[t | 4][i]);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (!(t & 8))
#pragma unroll
                for (int i = 0; i < 4; ++i) swap_l5(e[t][i], e[t | 8][i]);
        // now (g,i) = j1, tile = k0
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(e[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {re[i], im[i]};
                const u32 row = (u32)t + 16u * (4u * g + i);
                if constexpr (EPI) v = cmul(v, epi_twiddle(epi, out_pos(row), sg));
                buf[blk_base + (size_t)row * s] = pack(v);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// k_cols_small<R>: strided radix-R pass (R = 2, 4, 8), FP32 VALU, four adjacent
// columns per lane (16-byte loads).  Rows are in natural order within the pass.
template <int R, bool EPI>
__global__ __launch_bounds__(256) void k_cols_small(u32* __restrict__ buf, u32 s_log2, u32 n_items, int sign,
                                                    EpiParams epi) {
    const u32 item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= n_items) return;
    const u32 s = 1u << s_log2;
    const u32 quads = s >> 2;
    const u32 blk = item / quads;
    const u32 col = (item - blk * quads) * 4u;
    const size_t base = (size_t)blk * R * s + col;
    const float sg = sign < 0 ? 1.f : -1.f;
    u32 w[R][4];
#pragma unroll
    for (int j = 0; j < R; ++j) load4(w[j], buf + base + (size_t)j * s);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cf v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = unpack(w[j][q]);
        small_dft<R>(v, sign);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            cf o = v[j];
            if constexpr (EPI) o = cmul(o, epi_twiddle(epi, (u32)(base + (size_t)j * s + q), sg));
            w[j][q] = pack(o);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) store4(buf + base + (size_t)j * s, w[j]);
}

// ----------------------------------------------------------------------------
// Host-side launchers (called by the pass driver).
template <int T>
static hipError_t launch_rows_t(u32* buf, u32 n_tiles, int sign, const float2* w4096, const EpiParams* epi,
                                hipStream_t st) {
    const u32 waves = (n_tiles + 15) / 16;
    const u32 blocks = (waves + 3) / 4;
    if (epi) hipLaunchKernelGGL((k_rows<T, true>), dim3(blocks), dim3(256), 0, st, buf, n_tiles, sign, w4096, *epi);
    else hipLaunchKernelGGL((k_rows<T, false>), dim3(blocks), dim3(256), 0, st, buf, n_tiles, sign, w4096, EpiParams{});
    return hipGetLastError();
}

hipError_t launch_rows(int T, u32* buf, u32 n_tiles, int sign, const float2* w4096, const EpiParams* epi,
                       hipStream_t st) {
    switch (T) {
    case 1: return launch_rows_t<1>(buf, n_tiles, sign, w4096, epi, st);
    case 2: return launch_rows_t<2>(buf, n_tiles, sign, w4096, epi, st);
    case 4: return launch_rows_t<4>(buf, n_tiles, sign, w4096, epi, st);
    case 8: return launch_rows_t<8>(buf, n_tiles, sign, w4096, epi, st);
    case 16: return launch_rows_t<16>(buf, n_tiles, sign, w4096, epi, st);
    default: return hipErrorInvalidValue;
    }
}

template <int T>
static hipError_t launch_cols_t(u32* buf, u32 s_log2, u32 outer, int sign, const float2* w4096,
                                const EpiParams* epi, hipStream_t st) {
    const u32 waves = outer * ((1u << s_log2) >> 4);
    const u32 blocks = (waves + 3) / 4;
    if (epi) hipLaunchKernelGGL((k_cols<T, true>), dim3(blocks), dim3(256), 0, st, buf, s_log2, waves, sign, w4096, *epi);
    else hipLaunchKernelGGL((k_cols<T, false>), dim3(blocks), dim3(256), 0, st, buf, s_log2, waves, sign, w4096, EpiParams{});
    return hipGetLastError();
}

template <int R>
static hipError_t launch_cols_small_t(u32* buf, u32 s_log2, u32 outer, int sign, const EpiParams* epi,
                                      hipStream_t st) {
    const u32 items = outer * ((1u << s_log2) >> 2);
    const u32 blocks = (items + 255) / 256;
    if (epi) hipLaunchKernelGGL((k_cols_small<R, true>), dim3(blocks), dim3(256), 0, st, buf, s_log2, items, sign, *epi);
    else hipLaunchKernelGGL((k_cols_small<R, false>), dim3(blocks), dim3(256), 0, st, buf, s_log2, items, sign, EpiParams{});
    return hipGetLastError();
}

hipError_t launch_cols(int R, u32* buf, u32 s_log2, u32 outer, int sign, const float2* w4096, const EpiParams* epi,
                       hipStream_t st) {
    switch (R) {
    case 2: return launch_cols_small_t<2>(buf, s_log2, outer, sign, epi, st);
    case 4: return launch_cols_small_t<4>(buf, s_log2, outer, sign, epi, st);
    case 8: return launch_cols_small_t<8>(buf, s_log2, outer, sign, epi, st);
    case 16: return launch_cols_t<1>(buf, s_log2, outer, sign, w4096, epi, st);
    case 32: return launch_cols_t<2>(buf, s_log2, outer, sign, w4096, epi, st);
    case 64: return launch_cols_t<4>(buf, s_log2, outer, sign, w4096, epi, st);
    case 128: return launch_cols_t<8>(buf, s_log2, outer, sign, w4096, epi, st);
    case 256: return launch_cols_t<16>(buf, s_log2, outer, sign, w4096, epi, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace tcfft
