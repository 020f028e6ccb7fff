// fft_blocks.hpp -- device building blocks shared by the pass kernels
// (fft_kernels.hip) and the fused convolution kernels (conv.hip): epilogue
// twiddles, the in-wave 4096-point DIT, and the 4096*W-point single-pass
// transform (load, per-line DIT, LDS exchange, radix-W stage, store).
#pragma once
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace tcfft {

// Start delay of a first-pass workgroup under schedule s (see StartSched):
// s.delay sleep quanta of 1024 cycles per group index, for the workgroups of
// the first resident wave only.  The loop count is uniform and bounded.
__device__ __forceinline__ void start_delay(const StartSched& s) {
    if (s.delay && blockIdx.x < s.resident) {
        const u32 quanta = ((blockIdx.x >> 3) % s.groups) * s.delay;
        for (u32 i = 0; i < quanta; ++i) __builtin_amdgcn_s_sleep(16);
    }
}

// Epilogue twiddle for the next pass: element at position P (within the
// transform axis, P = global_pos >> shift) gets W_{s*R}^{c*rev(j)} with
// c = P mod s, j = (P / s) mod R and rev the two-digit DIT reversal of the
// next pass (radix ra first, then R/ra).  W_M^m = hi[m >> lo_bits]*lo[m & lo_mask].
// Digit reversal of a row index j (0 <= j < R) of the next pass, whose DIT
// digits are [2^ra, 2^rb, 2^rc] (rc = r - ra - rb, often 0):
// j = j0 + 2^ra*j1 + 2^(ra+rb)*j2 -> j2 + 2^rc*j1 + 2^(rc+rb)*j0.
__device__ __forceinline__ u32 epi_rev_row(const EpiParams& e, u32 j) {
    const u32 rc = e.r_log2 - e.ra_log2 - e.rb_log2;
    const u32 j0 = j & ((1u << e.ra_log2) - 1u);
    const u32 j1 = (j >> e.ra_log2) & ((1u << e.rb_log2) - 1u);
    const u32 j2 = j >> (e.ra_log2 + e.rb_log2);
    return j2 | (j1 << rc) | (j0 << (rc + e.rb_log2));
}

__device__ __forceinline__ cf epi_twiddle(const EpiParams& e, u32 gpos, float sgn) {
    const u32 P = gpos >> e.shift;
    const u32 c = P & ((1u << e.s_log2) - 1u);
    const u32 j = (P >> e.s_log2) & ((1u << e.r_log2) - 1u);
    const u32 rj = epi_rev_row(e, j);
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
    return epi_rev_row(e, blk & ((1u << e.r_log2) - 1u));
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
    // Stages 0 and 1 only combine points of one tile, so both run tile by
    // tile: tile t needs only its own load, and with the loads issued in tile
    // order (they complete in issue order) its work starts as soon as that
    // load has arrived instead of after the last one.
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        // stage 0: DFT over d0 -> lane%16 = k0, (g,i) = d1
        f32x4 re, im;
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(cmul(cf{re[i], im[i]}, tw1[i]));
        // stage 1: DFT over d1 -> lane%16 = k1, (g,i) = k0, tile = d2; twiddle W_4096^{c*d2}
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

// Twiddle-table values of the per-line phase, loaded before the data (global
// loads complete in issue order, so values loaded after the data would hold
// the first stage until the whole signal has arrived).
//   Phase-1 twiddle W_N^{w*c} = W_Q^{c*e} with root W_Q and exponent e:
//     no epilogue: Q = N, e = w;
//     epilogue (W = 16, next pass of radix R = 2^r at step N over chunks
//     j = sig mod R): the next pass's twiddle W_{NR}^{(c + 4096k)*rev(j)}
//     splits into W_{NR}^{c*rev(j)}, merged here as Q = N*R,
//     e = w*R + rev(j), and W_{NR}^{4096*k*rev(j)}, a row scale of the
//     final-stage operand (output frequency k).
//   rec0: W_Q^{e*(k0 + 16*k1)} (k0 = 4g + i, k1 = lane%16), constant over
//         k2 for c = k0 + 16*k1 + 256*k2, merged into the stage-1 recurrence;
//   s2:   W_Q^{e*256*k2} as a row scale of the stage-2 operand (f = k2);
//   s3:   the final-stage row scale under the epilogue.
struct RawQ {
    float2 h, l;  // W_Q^m = h * l (two-level table entries)
};

struct BigTw {
    Dit4096Tw tw;
    RawQ rec0[4], s2, s3;
};

template <int W, bool EPI>
__device__ __forceinline__ void big_tw_load(BigTw& bt, int lane, int w, u32 sig, const float2* __restrict__ w4096,
                                            const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo,
                                            u32 wn_lo_bits, const EpiParams& epi) {
    static_assert(!EPI || W == 16, "the epilogue twiddle is folded for W = 16 only");
    const int c16 = lane & 15, g = lane >> 4;
    u32 e_w = (u32)w, rj = 0u;
    if constexpr (EPI) {
        rj = epi_rev(epi, sig);
        e_w = ((u32)w << epi.r_log2) + rj;
    }
    const float2* q_hi = EPI ? epi.hi : wn_hi;
    const float2* q_lo = EPI ? epi.lo : wn_lo;
    const u32 q_bits = EPI ? epi.lo_bits : wn_lo_bits;
    const u32 q_mask = EPI ? epi.m_mask : 4096u * W - 1u;
    auto wq_load = [&](u32 m) -> RawQ {
        m &= q_mask;
        return RawQ{q_hi[m >> q_bits], q_lo[m & ((1u << q_bits) - 1u)]};
    };
    dit4096_tw_load(bt.tw, lane, w4096);
#pragma unroll
    for (int i = 0; i < 4; ++i) bt.rec0[i] = wq_load(e_w * (4u * g + i + 16u * c16));
    bt.s2 = wq_load(256u * e_w * (u32)c16);
    bt.s3 = wq_load(4096u * rj * (u32)c16);
}

template <int SIGN>
__device__ __forceinline__ cf raw_q(const RawQ& r) {
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    return cmul(cf{r.h.x, sg * r.h.y}, cf{r.l.x, sg * r.l.y});
}

// Phase 1: the in-wave 4096-point DIT of this wave's line d (the digit-
// reversed chunk 4096*w .. 4096*w + 4095 of the signal, loaded as tile t =
// d2, lane (c16 = d1, g), word j = d0), times W_N^{w*c}; the result is left
// in e with (g,i) = k0, tile = k2, lane%16 = k1, c = k0 + 16*k1 + 256*k2.
// With scale != 1 every result is also multiplied by scale, at no cost: the
// factor rides on the start of the stage-1 recurrence, which multiplies
// every point of the line once (in FP32, before the rounding to fp16).
template <int SIGN>
__device__ __forceinline__ void big_phase1(u32 (&d)[16][4], u32 (&e)[16][4], const BigTw& bt, const Dft16Op& op,
                                           float scale = 1.f) {
    cf rec0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        rec0[i] = raw_q<SIGN>(bt.rec0[i]);
        rec0[i] = cf{rec0[i].x * scale, rec0[i].y * scale};
    }
    const cf s2 = raw_q<SIGN>(bt.s2);
    auto make_op2 = [&] { return make_dft16_scaled(lane_id(), SIGN, s2); };
    dit4096<true>(d, SIGN, bt.tw, op, make_op2, rec0, [](int, int, cf v) -> cf { return v; });
    transpose_gi_tile(d, e);
}

// Phase 2: LDS exchange of the W lines (layout [w][sigma(c)]) and the
// radix-W stage over w, storing X[c + 4096*k3] at sbase (4 adjacent columns
// per lane).  With SCALE the FP32 results are multiplied by `scale` first.
// The stores are non-temporal, or with CACHED cached, for a pass whose
// output the next pass reads back from the MALL (a template parameter: the
// compiler merges a runtime choice between the two into one cached store).
// Every barrier is a workgroup barrier: all groups of a workgroup run the
// same sequence.
template <int W, int ROUNDS, bool EPI, int SIGN, bool SCALE = false, bool CACHED = false>
__device__ __forceinline__ void big_phase2(u32 (&e)[16][4], int w, u32* lds, u32* __restrict__ sbase,
                                           const BigTw& bt, const Dft16Op& op, bool active, float scale = 1.f) {
    static_assert(!SCALE || W < 16, "output scaling is implemented for the VALU radix-W stage");
    constexpr int sign = SIGN;
    using C = BigCfg<W, ROUNDS>;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
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
            if (r == 0) op3 = make_dft16_scaled(lane_id(), sign, raw_q<SIGN>(bt.s3));
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
                    if (active) {
                        if constexpr (CACHED)
                            store4(sbase + c0 + 4096u * (4u * g + i), q);
                        else
                            store4_nt(sbase + c0 + 4096u * (4u * g + i), q);
                    }
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
                    if constexpr (W == 2 && !SCALE) {
                        // radix 2 on the packed words: one packed fp16 add
                        // and subtract, each correctly rounded
                        o[0][j] = hadd2(v[0][j], v[1][j]);
                        o[1][j] = hsub2(v[0][j], v[1][j]);
                    } else {
                        cf x[W];
#pragma unroll
                        for (int s = 0; s < W; ++s) x[s] = unpack(v[s][j]);
                        small_dft<W>(x, sign);
#pragma unroll
                        for (int s = 0; s < W; ++s)
                            o[s][j] = SCALE ? pack(x[s].x * scale, x[s].y * scale) : pack(x[s]);
                    }
                }
                const u32 c0 = 2048u * r + 4u * lane + 256u * (q + QN * w);
#pragma unroll
                for (int s = 0; s < W; ++s)
                    if (active) {
                        if constexpr (CACHED)
                            store4(sbase + c0 + 4096u * s, o[s]);
                        else
                            store4_nt(sbase + c0 + 4096u * s, o[s]);
                    }
            }
        }
    }
}

// big_phase2 for W = 2, 4, 8 that keeps the results in registers instead of
// storing them: o[(r*QN + q)*W + s] holds the four words of positions
// 2048r + 4*lane + 256*(q + QN*w) + 4096*s + (0..3) (the words big_phase2
// stores there).  The caller synchronises before reusing the LDS.
template <int W, int ROUNDS, int SIGN>
__device__ __forceinline__ void big_phase2_regs(u32 (&e)[16][4], int w, u32* lds, u32 (&o)[16][4]) {
    static_assert(W >= 2 && W <= 8, "VALU radix-W stage");
    using C = BigCfg<W, ROUNDS>;
    constexpr int QN = C::COLS / (256 * W);
    static_assert(ROUNDS * QN * W == 16, "one slot per output group");
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < C::ROUNDS; ++r) {
        constexpr int TPR = 16 / C::ROUNDS;
        if (r) __syncthreads();
#pragma unroll
        for (int tt = 0; tt < TPR; ++tt) {
            const u32 off = (u32)w * C::COLS + lds_sigma(4u * g + 16u * c16 + 256u * tt);
            store4(lds + off, e[r * TPR + tt]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            u32 v[W][4];
#pragma unroll
            for (int s = 0; s < W; ++s)
                load4(v[s], lds + (u32)s * C::COLS + lds_sigma(4u * lane + 256u * (q + QN * w)));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                cf x[W];
#pragma unroll
                for (int s = 0; s < W; ++s) x[s] = unpack(v[s][j]);
                small_dft<W>(x, SIGN);
#pragma unroll
                for (int s = 0; s < W; ++s) o[(r * QN + q) * W + s][j] = pack(x[s]);
            }
        }
    }
}

// One signal of N = 4096*W at sbase, processed by the W waves of a group
// (this wave is number w) with lds = the group's W*COLS words of LDS.  sig
// is the signal's index (epilogue only).
template <int W, int ROUNDS, bool EPI, int SIGN, bool CACHED = false>
__device__ __forceinline__ void big_body(u32* __restrict__ sbase, int w, u32* lds, u32 sig,
                                         const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                         const float2* __restrict__ wn_lo, u32 wn_lo_bits, const EpiParams& epi,
                                         bool active = true) {
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    BigTw bt;
    big_tw_load<W, EPI>(bt, lane, w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, epi);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
    {
        const u32* src = sbase + 4096 * w + 16 * c16 + 4 * g;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            load4_nt(d[t], src + 256 * t);
            __builtin_amdgcn_sched_barrier(0);  // keep the tile order of the loads
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    const Dft16Op op = make_dft16(lane, SIGN);
    u32 e[16][4];
    big_phase1<SIGN>(d, e, bt, op);
    big_phase2<W, ROUNDS, EPI, SIGN, false, CACHED>(e, w, lds, sbase, bt, op, active);
}

}  // namespace tcfft
