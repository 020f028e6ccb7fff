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

#include <mutex>

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

// Strided passes see their output at (block blk, row, column col) of a
// buffer viewed as [blk][R][s]; along the axis that is P = blk*R*sa + row*sa
// + xc with sa = s >> shift and xc = col >> shift, so the next pass's twiddle
// is W_M^{(row*sa + xc)*rev(blk mod R')}: a per-column factor times a
// per-row factor, both powers of one root.  epi_rev gives rev(blk mod R'),
// epi_w gives W_M^m from the two-level table.
__device__ __forceinline__ u32 epi_rev(const EpiParams& e, u32 blk) {
    const u32 j = blk & ((1u << e.r_log2) - 1u);
    const u32 j0 = j & ((1u << e.ra_log2) - 1u);
    return (j >> e.ra_log2) | (j0 << (e.r_log2 - e.ra_log2));
}

__device__ __forceinline__ cf epi_w(const EpiParams& e, u32 m, float sgn) {
    m &= e.m_mask;
    return cmul(twiddle(e.hi, m >> e.lo_bits, sgn), twiddle(e.lo, m & ((1u << e.lo_bits) - 1u), sgn));
}

// ----------------------------------------------------------------------------
// In-wave 4096-point DIT over the digits d0, d1, d2 of a 4096-element chunk.
// In : d[t][j] = tile t (d2), lane (c16 = d1, g), word j (d0 = 4g + j).
// Out: d[t][i] = tile t (k0), lane%16 = k1, (g, i) = k2, after post(t, i, v)
//      is applied to every FP32 result of the last stage.
// Stage-1 twiddle W_256^{k0*d1} is a per-lane constant; the stage-2 twiddle
// W_4096^{(k0+16k1)*d2} is a per-lane recurrence over the tiles.  The
// last stage uses the operand make_op2() (B-form), which may carry a row
// scale; it is built after stage 1 so that it holds no registers before.
// With REC0 the stage-2 inputs of column c = k0 + 16*k1 are also multiplied
// by rec0[i] (k0 = 4g + i, k1 = lane%16): a per-column factor of the caller
// merged into the stage-1 recurrence (rec = rec0 * W_4096^{c*d2}).
// The per-lane twiddles come from the caller (Dit4096Tw), loaded before the
// data: global loads complete in issue order, so a table load issued after
// the data loads would hold stage 0 until the whole chunk has arrived.
struct Dit4096Tw {
    float2 tw1[4], step[4];  // raw table entries: W_256^{k0*d1}, W_4096^{k0+16*k1}
};

__device__ __forceinline__ void dit4096_tw_load(Dit4096Tw& tw, int lane, const float2* __restrict__ w4096) {
    const int c16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tw.tw1[i] = w4096[((c16 * (4 * g + i)) & 255) * 16];
        tw.step[i] = w4096[16 * c16 + 4 * g + i];
    }
}

template <bool REC0, class MakeOp2, class Post>
__device__ __forceinline__ void dit4096(u32 (&d)[16][4], int sign, const Dit4096Tw& tw, const Dft16Op& op,
                                        MakeOp2 make_op2, const cf (&rec0)[4], Post post) {
    const float sg = sign < 0 ? 1.f : -1.f;
    cf tw1[4], step[4], rec[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tw1[i] = cf{tw.tw1[i].x, sg * tw.tw1[i].y};
        step[i] = cf{tw.step[i].x, sg * tw.step[i].y};
        rec[i] = REC0 ? rec0[i] : step[i];
    }
    // stage 0: DFT over d0 -> lane%16 = k0, (g,i) = d1
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(cmul(cf{re[i], im[i]}, tw1[i]));
    }
    // stage 1: DFT over d1 -> lane%16 = k1, (g,i) = k0, tile = d2; twiddle W_4096^{c*d2}
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf v = {re[i], im[i]};
            if (REC0 || t) {
                v = cmul(v, rec[i]);
                if (t < 15) rec[i] = cmul(rec[i], step[i]);
            }
            d[t][i] = pack(v);
        }
    }
    u32 e[16][4];
    transpose_gi_tile(d, e);  // (g,i) = d2, tile = k0
    const Dft16Op op2 = make_op2();
    // stage 2 (B-form): columns k1, points d2 -> (g,i) = k2
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_b(e[t], op2, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(post(t, i, cf{re[i], im[i]}));
    }
}

// ----------------------------------------------------------------------------
// k_rows<T>: lines of L = 256*T contiguous elements, digit-reversed for the
// radix list [16, 16] (+ [T] when T > 1).  Position in a line:
// p = d0 + 16*d1 + 256*d2.  Wave layout at load: tile t = 256-element block,
// lane (c16 = d1, g) holds d0 = 4g..4g+3 in its four words (one 16-byte load).
// One call's work: NT tiles of 256 elements (NT = 16, or a multiple of T
// below 16 to bound register use); tile t lives at tile_ptr(t) (already
// offset by this lane's 16*c16 + 4*g); tiles t >= ntl are absent.  tile0 is
// the global tile index of tile 0 (used by the epilogue only).
template <int T, bool EPI, int SIGN, int NT = 16, class TilePtr>
__device__ __forceinline__ void rows_body(TilePtr tile_ptr, u32 ntl, u32 tile0, const float2* __restrict__ w4096,
                                          const EpiParams& epi) {
    static_assert(NT % T == 0 && NT <= 16 && (T < 16 || NT == 16), "tile count per call");
    constexpr int sign = SIGN;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const float sg = sign < 0 ? 1.f : -1.f;

    // table loads first, then the data (see Dit4096Tw)
    Dit4096Tw tw;
    dit4096_tw_load(tw, lane, w4096);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        if ((u32)t < ntl) load4_nt(d[t], tile_ptr(t));
        else d[t][0] = d[t][1] = d[t][2] = d[t][3] = 0u;
    }
    __builtin_amdgcn_sched_barrier(0);
    const Dft16Op op = make_dft16(lane, sign);
    cf tw1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) tw1[i] = cf{tw.tw1[i].x, sg * tw.tw1[i].y};

    // Stage 0 (T < 16; dit4096 does its own): DFT over d0 (A-form) ->
    // lane%16 = k0, rows (g,i) = d1.  Twiddle W_256^{k0*d1} is a per-lane constant.
    if constexpr (T < 16) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 re, im;
            dft16_a(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) d[t][i] = pack(cmul(cf{re[i], im[i]}, tw1[i]));
        }
    }

    if constexpr (T == 1) {
        // Stage 1: DFT over d1 -> lane%16 = k1, rows (g,i) = k0; natural
        // position c = k0 + 16*k1 = 4g + i + 16*c16: four adjacent words.
#pragma unroll
        for (int t = 0; t < NT; ++t) {
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
            if ((u32)t < ntl) store4(tile_ptr(t), d[t]);
        }
    } else if constexpr (T < 16) {
        // Stage 1 on each tile of a line, then radix-T across the line's tiles:
        // X[c + 256*k2] = sum_d2 W_{256T}^{c*d2} Y_d2[c] W_T^{d2*k2}.
        constexpr int STRIDE = 16 / T;  // W_{256T}^m = W_4096^{m*STRIDE}
#pragma unroll
        for (int gi = 0; gi < NT / T; ++gi) {
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
                if ((u32)t < ntl) store4(tile_ptr(t), d[t]);
            }
        }
    } else {
        // T == 16: one 4096-point line per wave (shared in-wave DIT).
        const cf none[4] = {};
        dit4096<false>(d, sign, tw, op, [&] { return op; }, none, [&](int t, int i, cf v) -> cf {
            if constexpr (EPI) {
                const u32 k = (u32)t + 16u * c16 + 256u * (4u * g + i);
                v = cmul(v, epi_twiddle(epi, tile0 * 256u + k, sg));
            }
            return v;
        });
        // (g,i) = k2, tile = k0 -> (g,i) = k0, tile = k2: four adjacent words per lane
        u32 e[16][4];
        transpose_gi_tile(d, e);
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if ((u32)t < ntl) store4_nt(tile_ptr(t), e[t]);
    }
}

template <int T, bool EPI, int SIGN>
__global__ __launch_bounds__(256) void k_rows(u32* __restrict__ buf, u32 n_tiles,
                                              const float2* __restrict__ w4096, EpiParams epi) {
    const int lane = threadIdx.x & 63;
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const u32 tile0 = wave * 16u;
    if (tile0 >= n_tiles) return;
    u32* const base = buf + (size_t)tile0 * 256 + 16 * (lane & 15) + 4 * (lane >> 4);
    rows_body<T, EPI, SIGN>([&](int t) { return base + t * 256; }, min(16u, n_tiles - tile0), tile0, w4096, epi);
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

    // Epilogue: output row k0 + 16*t (T < 16), t + 16*k0 (T = 16) or k0
    // (T = 1), k0 = 4g + i, gets F0[i] * S^t with F0[i] = W_M^{(xc + k0*sa*u)*rj}
    // (u = 16 for T = 16, else 1) and S = W_M^{sa*rj*(T == 16 ? 1 : 16)}.
    cf F0[4], S = {1.f, 0.f};
    if constexpr (EPI) {
        const u32 rj = epi_rev(epi, blk);
        const u32 sa = s >> epi.shift, xc = (col0 + (u32)c16) >> epi.shift;
        const u32 u = T == 16 ? 16u : 1u;
#pragma unroll
        for (int i = 0; i < 4; ++i) F0[i] = epi_w(epi, (xc + (4u * g + i) * u * sa) * rj, sg);
        if constexpr (T > 1) S = epi_w(epi, sa * rj * (T == 16 ? 1u : 16u), sg);
    }

    if constexpr (T == 1) {
        f32x4 re, im;
        dft16_b(d[0], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf v = {re[i], im[i]};
            const u32 row = 4u * g + i;
            if constexpr (EPI) v = cmul(v, F0[i]);
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
            cf f = EPI ? F0[i] : cf{1.f, 0.f};
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const u32 row = k0 + 16u * t;
                cf o = v[t];
                if constexpr (EPI) {
                    o = cmul(o, f);
                    f = cmul(f, S);
                }
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
                if constexpr (EPI) {
                    v = cmul(v, F0[i]);
                    F0[i] = cmul(F0[i], S);
                }
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
    // Epilogue: output (row j, column col + q) gets W_M^{(j*sa + xc_q)*rj};
    // xc_q = xc + q when shift = 0, else xc (4 adjacent columns, one x).
    cf F[R], Q = {1.f, 0.f};
    if constexpr (EPI) {
        const u32 rj = epi_rev(epi, blk);
        const u32 sa = s >> epi.shift, xc = col >> epi.shift;
#pragma unroll
        for (int j = 0; j < R; ++j) F[j] = epi_w(epi, (xc + (u32)j * sa) * rj, sg);
        if (epi.shift == 0) Q = epi_w(epi, rj, sg);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cf v[R];
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = unpack(w[j][q]);
        small_dft<R>(v, sign);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            cf o = v[j];
            if constexpr (EPI) {
                o = cmul(o, F[j]);
                F[j] = cmul(F[j], Q);
            }
            w[j][q] = pack(o);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) store4(buf + base + (size_t)j * s, w[j]);
}

// ----------------------------------------------------------------------------
// k_big<W>: one signal of N = 4096*W (W = 2, 4, 8, 16) per workgroup of W
// waves, in a single HBM round trip.  Radix list [16, 16, 16, W]; position
// p = d0 + 16*d1 + 256*d2 + 4096*w, wave w loads chunk w.
//  phase 1: each wave does the 4096-point DIT of its chunk in registers and
//           multiplies by the final-stage twiddle W_N^{w*c}, c = k0+16k1+256k2:
//           W_N^{256*w*k2} is folded into the stage-2 operand (row scale),
//           W_N^{w*(k0+16k1)} is a per-lane recurrence over the tiles.
//  phase 2: LDS exchange (layout [w][sigma(c)], b128 both ways) and the
//           radix-W stage over w (MFMA for W = 16, VALU for W = 2, 4, 8);
//           output position c + 4096*k3.  Two rounds of 2048 columns for
//           W >= 8 keep the exchange at W*8 KiB of LDS.
// sigma(c) = c ^ ((c >> 4) & 15) << 2 is a bijection that keeps 16-lane
// groups conflict-free both for writers (lane%16 = c bits 4-7) and for
// readers (lane%16 = c bits 2-5); readers own 4 adjacent columns per lane so
// the final global stores are 256-byte runs per 16 lanes.
__device__ __forceinline__ u32 lds_sigma(u32 c) { return c ^ (((c >> 4) & 15u) << 2); }

template <int W, int ROUNDS_ = (W >= 8 ? 2 : 1)>
struct BigCfg {
    static constexpr int ROUNDS = ROUNDS_;
    static constexpr int COLS = 4096 / ROUNDS;  // columns exchanged per round
    static constexpr int LDS_WORDS = W * COLS;
    static constexpr u32 N = 4096u * W;
};

// One signal of N = 4096*W at sbase, processed by the W waves of a group
// (this wave is number w) with lds = the group's W*COLS words of LDS.
// Every barrier is a workgroup barrier: all groups of a workgroup run the
// same sequence.  sig is the signal's index (epilogue positions only).
template <int W, int ROUNDS, bool EPI, int SIGN>
__device__ __forceinline__ void big_body(u32* __restrict__ sbase, int w, u32* lds, u32 sig,
                                         const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                         const float2* __restrict__ wn_lo, u32 wn_lo_bits, const EpiParams& epi,
                                         bool active = true) {
    static_assert(!EPI || W == 16, "the epilogue twiddle is folded for W = 16 only");
    constexpr int sign = SIGN;
    using C = BigCfg<W, ROUNDS>;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const float sg = sign < 0 ? 1.f : -1.f;

    // Phase-1 twiddle W_N^{w*c} = W_Q^{c*e} with root W_Q and exponent e:
    //   no epilogue: Q = N, e = w;
    //   epilogue (W = 16, next pass of radix R = 2^r at step N over chunks
    //   j = sig mod R): the next pass's twiddle W_{NR}^{(c + 4096k)*rev(j)}
    //   splits into W_{NR}^{c*rev(j)}, merged here as Q = N*R,
    //   e = w*R + rev(j), and W_{NR}^{4096*k*rev(j)}, a row scale of the
    //   final-stage operand (output frequency k).
    u32 e_w = (u32)w, rj = 0u;
    if constexpr (EPI) {
        const u32 j = sig & ((1u << epi.r_log2) - 1u);
        const u32 j0 = j & ((1u << epi.ra_log2) - 1u);
        rj = (j >> epi.ra_log2) | (j0 << (epi.r_log2 - epi.ra_log2));
        e_w = ((u32)w << epi.r_log2) + rj;
    }
    // W_Q^m = hi[m >> lo_bits] * lo[m & (2^lo_bits - 1)] (two-level table of Q)
    const float2* q_hi = EPI ? epi.hi : wn_hi;
    const float2* q_lo = EPI ? epi.lo : wn_lo;
    const u32 q_bits = EPI ? epi.lo_bits : wn_lo_bits;
    const u32 q_mask = EPI ? epi.m_mask : C::N - 1u;
    struct RawQ {
        float2 h, l;
    };
    auto wq_load = [&](u32 m) -> RawQ {
        m &= q_mask;
        return RawQ{q_hi[m >> q_bits], q_lo[m & ((1u << q_bits) - 1u)]};
    };
    auto wq_val = [&](const RawQ& r) -> cf { return cmul(cf{r.h.x, sg * r.h.y}, cf{r.l.x, sg * r.l.y}); };

    // Table loads first, then the data: global loads complete in issue
    // order, so table values loaded after the data would hold the first
    // stage until the whole signal has arrived.
    //   rec0: W_Q^{e*(k0 + 16*k1)} (k0 = 4g + i, k1 = lane%16), constant over
    //         k2 for c = k0 + 16*k1 + 256*k2, merged into the stage-1 recurrence;
    //   s2:   W_Q^{e*256*k2} as a row scale of the stage-2 operand (f = k2);
    //   s3:   the final-stage row scale under the epilogue.
    Dit4096Tw tw;
    dit4096_tw_load(tw, lane, w4096);
    RawQ r_rec0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r_rec0[i] = wq_load(e_w * (4u * g + i + 16u * c16));
    const RawQ r_s2 = wq_load(256u * e_w * (u32)c16);
    const RawQ r_s3 = wq_load(4096u * rj * (u32)c16);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
    {
        const u32* src = sbase + 4096 * w + 16 * c16 + 4 * g;
#pragma unroll
        for (int t = 0; t < 16; ++t) load4_nt(d[t], src + 256 * t);
    }
    __builtin_amdgcn_sched_barrier(0);

    const Dft16Op op = make_dft16(lane, sign);
    cf rec0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) rec0[i] = wq_val(r_rec0[i]);
    const cf s2 = wq_val(r_s2);
    auto make_op2 = [&] { return make_dft16_scaled(lane_id(), sign, s2); };
    dit4096<true>(d, sign, tw, op, make_op2, rec0, [](int, int, cf v) -> cf { return v; });
    // d: tile = k0, lane%16 = k1, (g,i) = k2.  Writer layout: word = c bits 0-1
    u32 e[16][4];
    transpose_gi_tile(d, e);  // (g,i) = k0, tile = k2: c = (4g+i) + 16*c16 + 256*t
    // final-stage operand (W = 16): rows k scaled by W_Q^{4096*k*rev(j)} under
    // the epilogue; built after the first exchange barrier (a fresh lane id)
    // so that it does not hold registers during the in-wave DIT
    Dft16Op op3 = op;

#pragma unroll
    for (int r = 0; r < C::ROUNDS; ++r) {
        constexpr int TPR = 16 / C::ROUNDS;  // tiles per round
        if (r) __syncthreads();              // previous round's reads are done
#pragma unroll
        for (int tt = 0; tt < TPR; ++tt) {
            const u32 off = (u32)w * C::COLS + lds_sigma(4u * g + 16u * c16 + 256u * tt);
            store4(lds + off, e[r * TPR + tt]);
        }
        __syncthreads();
        if constexpr (EPI) {
            if (r == 0) op3 = make_dft16_scaled(lane_id(), sign, wq_val(r_s3));
        }
        if constexpr (W == 16) {
            // reader wave w': columns c = j + 4*c16 + 64h + 128w' (+2048r)
            u32 v[2][4][4];  // [h][i = point (w = 4g+i)][j]
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const u32 off = (u32)(4 * g + i) * C::COLS + lds_sigma(4u * c16 + 64u * h + 128u * w);
                    load4(v[h][i], lds + off);
                }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                u32 o[4][4];  // [j][i' = k3 word]
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    u32 a[4] = {v[h][0][j], v[h][1][j], v[h][2][j], v[h][3][j]};
                    f32x4 re, im;
                    dft16_b(a, op3, re, im);
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[j][i] = pack(cf{re[i], im[i]});
                }
                const u32 c0 = 2048u * r + 4u * c16 + 64u * h + 128u * w;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    u32 q[4] = {o[0][i], o[1][i], o[2][i], o[3][i]};
                    if (active) store4_nt(sbase + c0 + 4096u * (4u * g + i), q);
                }
            }
        } else {
            // reader wave w': columns c = j + 4*lane + 256*(q + QN*w') (+2048r)
            constexpr int QN = C::COLS / (256 * W);
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                u32 v[W][4];
#pragma unroll
                for (int s = 0; s < W; ++s)
                    load4(v[s], lds + (u32)s * C::COLS + lds_sigma(4u * lane + 256u * (q + QN * w)));
                u32 o[W][4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cf x[W];
#pragma unroll
                    for (int s = 0; s < W; ++s) x[s] = unpack(v[s][j]);
                    small_dft<W>(x, sign);
#pragma unroll
                    for (int s = 0; s < W; ++s) o[s][j] = pack(x[s]);
                }
                const u32 c0 = 2048u * r + 4u * lane + 256u * (q + QN * w);
#pragma unroll
                for (int s = 0; s < W; ++s)
                    if (active) store4_nt(sbase + c0 + 4096u * s, o[s]);
            }
        }
    }
}

template <int W, bool EPI, int SIGN>
__global__ __launch_bounds__(64 * W) void k_big(u32* __restrict__ buf, const float2* __restrict__ w4096,
                                                const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo,
                                                u32 wn_lo_bits, EpiParams epi) {
    extern __shared__ u32 lds[];
    big_body<W, BigCfg<W>::ROUNDS, EPI, SIGN>(buf + (size_t)blockIdx.x * BigCfg<W>::N, threadIdx.x >> 6, lds,
                                             blockIdx.x, w4096, wn_hi, wn_lo, wn_lo_bits, epi);
}

// ----------------------------------------------------------------------------
// Host-side launchers (called by the pass driver).
template <int T, int SIGN>
static hipError_t launch_rows_t(u32* buf, u32 n_tiles, const float2* w4096, const EpiParams* epi, hipStream_t st) {
    const u32 waves = (n_tiles + 15) / 16;
    const u32 blocks = (waves + 3) / 4;
    if (epi) hipLaunchKernelGGL((k_rows<T, true, SIGN>), dim3(blocks), dim3(256), 0, st, buf, n_tiles, w4096, *epi);
    else hipLaunchKernelGGL((k_rows<T, false, SIGN>), dim3(blocks), dim3(256), 0, st, buf, n_tiles, w4096, EpiParams{});
    return hipGetLastError();
}

template <int T>
static hipError_t launch_rows_t(u32* buf, u32 n_tiles, int sign, const float2* w4096, const EpiParams* epi,
                                hipStream_t st) {
    return sign < 0 ? launch_rows_t<T, -1>(buf, n_tiles, w4096, epi, st)
                    : launch_rows_t<T, +1>(buf, n_tiles, w4096, epi, st);
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

// Raise the dynamic-LDS limit of a kernel once per process (thread-safe).
template <class K>
static hipError_t set_lds_once(K kernel, size_t bytes) {
    static std::once_flag flag;
    static hipError_t status = hipSuccess;
    std::call_once(flag, [&] {
        status = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    });
    return status;
}

template <int W, int SIGN>
static hipError_t launch_big_t(u32* buf, u32 n_signals, const float2* w4096, const float2* wn_hi,
                               const float2* wn_lo, u32 lo_bits, const EpiParams* epi, hipStream_t st) {
    const size_t lds = BigCfg<W>::LDS_WORDS * sizeof(u32);
    if (epi) {
        // the folded epilogue covers a first pass of 65536 contiguous points
        // followed by a pass at step 65536 along the same axis
        if constexpr (W == 16) {
            if (epi->shift != 0 || epi->s_log2 != 16) return hipErrorInvalidValue;
            auto k = k_big<W, true, SIGN>;
            hipError_t e = set_lds_once(k, lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, buf, w4096, wn_hi, wn_lo, lo_bits, *epi);
        } else {
            return hipErrorInvalidValue;
        }
    } else {
        auto k = k_big<W, false, SIGN>;
        hipError_t e = set_lds_once(k, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, buf, w4096, wn_hi, wn_lo, lo_bits,
                           EpiParams{});
    }
    return hipGetLastError();
}

template <int W>
static hipError_t launch_big_t(u32* buf, u32 n_signals, int sign, const float2* w4096, const float2* wn_hi,
                               const float2* wn_lo, u32 lo_bits, const EpiParams* epi, hipStream_t st) {
    return sign < 0 ? launch_big_t<W, -1>(buf, n_signals, w4096, wn_hi, wn_lo, lo_bits, epi, st)
                    : launch_big_t<W, +1>(buf, n_signals, w4096, wn_hi, wn_lo, lo_bits, epi, st);
}

hipError_t launch_big(int W, u32* buf, u32 n_signals, int sign, const float2* w4096, const float2* wn_hi,
                      const float2* wn_lo, u32 lo_bits, const EpiParams* epi, hipStream_t st) {
    switch (W) {
    case 2: return launch_big_t<2>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, st);
    case 4: return launch_big_t<4>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, st);
    case 8: return launch_big_t<8>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, st);
    case 16: return launch_big_t<16>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, st);
    default: return hipErrorInvalidValue;
    }
}

// ----------------------------------------------------------------------------
// k_multi: many transforms of uneven power-of-two sizes (2^8 .. 2^16) in one
// persistent launch.  The host cuts the batch into tasks of one size class
// and at most 65536 elements: 16/T lines of L = 256*T per wave for L <= 4096
// (rows_body), or 16/W signals of N = 4096*W per workgroup for larger N
// (big_body, W waves and W*8 KiB of LDS each).  Workgroup b runs tasks
// b, b + grid, b + 2*grid, ... (no atomics, so the launch is graph-safe).
template <int T, int SIGN>
__device__ __forceinline__ void multi_rows(u32* buf, const MultiTask& tk, const u32* __restrict__ offsets,
                                           const float2* w4096) {
    constexpr int LPW = 16 / T;  // lines per wave
    // wave-uniform (scalar) so that the per-line offsets are scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int first_line = wave * LPW;
    const int lines = min((int)tk.count - first_line, LPW);
    if (lines <= 0) return;
    const u32 lane_off = 16u * (lane & 15) + 4u * (lane >> 4);
    const u32* offs = offsets + tk.first + first_line;
    if constexpr (T == 16) {
        rows_body<T, false, SIGN>([&](int t) { return buf + offs[t / T] + 256u * (t % T) + lane_off; },
                                  (u32)(lines * T), 0u, w4096, EpiParams{});
    } else {
        // calls of NT tiles keep the line data within the register budget of
        // a 1024-thread workgroup (128 VGPRs); NT per T is the largest that
        // compiles without spills
        constexpr int NT = (T == 1 || T == 8) ? 8 : 4;
        for (int t0 = 0; t0 < lines * T; t0 += NT) {
            const int ntl = lines * T - t0;
            rows_body<T, false, SIGN, NT>(
                [&](int t) { return buf + offs[(t0 + t) / T] + 256u * ((t0 + t) % T) + lane_off; }, (u32)ntl, 0u,
                w4096, EpiParams{});
        }
    }
}

template <int W, int SIGN>
__device__ __forceinline__ void multi_big(u32* buf, const MultiTask& tk, const u32* __restrict__ offsets,
                                          const float2* w4096, const MultiTables& tb, u32* lds) {
    constexpr int GROUPS = 16 / W;
    constexpr int M = W == 2 ? 13 : W == 4 ? 14 : W == 8 ? 15 : 16;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = wave / W;
    // A group without a signal (last task of a class) loads the task's first
    // signal, so that every load stays in bounds, and stores nothing.
    const bool active = grp < (int)tk.count;
    const u32 off = offsets[tk.first + (active ? grp : 0)];
    big_body<W, 2, false, SIGN>(buf + off, wave % W, lds + grp * (W * 2048), 0u, w4096, tb.hi[M - 13],
                                tb.lo[M - 13], (u32)(M / 2), EpiParams{}, active);
}

template <int SIGN>
__global__ __launch_bounds__(1024) void k_multi(u32* __restrict__ buf, const MultiTask* __restrict__ tasks,
                                                u32 n_tasks, const u32* __restrict__ offsets,
                                                const float2* __restrict__ w4096, MultiTables tb) {
    extern __shared__ u32 lds[];
    for (u32 ti = blockIdx.x; ti < n_tasks; ti += gridDim.x) {
        const MultiTask tk = tasks[ti];
        switch (tk.size_log2) {
        case 8: multi_rows<1, SIGN>(buf, tk, offsets, w4096); break;
        case 9: multi_rows<2, SIGN>(buf, tk, offsets, w4096); break;
        case 10: multi_rows<4, SIGN>(buf, tk, offsets, w4096); break;
        case 11: multi_rows<8, SIGN>(buf, tk, offsets, w4096); break;
        case 12: multi_rows<16, SIGN>(buf, tk, offsets, w4096); break;
        case 13: multi_big<2, SIGN>(buf, tk, offsets, w4096, tb, lds); break;
        case 14: multi_big<4, SIGN>(buf, tk, offsets, w4096, tb, lds); break;
        case 15: multi_big<8, SIGN>(buf, tk, offsets, w4096, tb, lds); break;
        case 16: multi_big<16, SIGN>(buf, tk, offsets, w4096, tb, lds); break;
        default: break;
        }
        __syncthreads();  // the next task reuses the LDS
    }
}

template <int SIGN>
static hipError_t launch_multi_t(u32* buf, const MultiTask* tasks, u32 n_tasks, const u32* offsets, u32 grid,
                                 const float2* w4096, const MultiTables& tb, hipStream_t st) {
    constexpr size_t lds = 16 * 2048 * sizeof(u32);  // 128 KiB
    auto k = k_multi<SIGN>;
    hipError_t e = set_lds_once(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1024), lds, st, buf, tasks, n_tasks, offsets, w4096, tb);
    return hipGetLastError();
}

hipError_t launch_multi(u32* buf, const MultiTask* tasks, u32 n_tasks, const u32* offsets, u32 grid, int sign,
                        const float2* w4096, const MultiTables& tb, hipStream_t st) {
    if (n_tasks == 0) return hipSuccess;
    if (grid == 0 || grid > n_tasks) grid = n_tasks;
    return sign < 0 ? launch_multi_t<-1>(buf, tasks, n_tasks, offsets, grid, w4096, tb, st)
                    : launch_multi_t<+1>(buf, tasks, n_tasks, offsets, grid, w4096, tb, st);
}

}  // namespace tcfft
