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
#include "fft_blocks.hpp"
#include "fft_kernels.hpp"

#include <map>
#include <tuple>
#include <type_traits>
#include <mutex>

namespace tcfft {

// ----------------------------------------------------------------------------
// k_rows<T>: lines of L = 256*T contiguous elements, digit-reversed for the
// radix list [16, 16] (+ [T] when T > 1).  Position in a line:
// p = d0 + 16*d1 + 256*d2.  Wave layout at load: tile t = 256-element block,
// lane (c16 = d1, g) holds d0 = 4g..4g+3 in its four words (one 16-byte load).
// One call's work: NT tiles of 256 elements (NT = 16, or a multiple of T
// below 16 to bound register use); tile t lives at tile_ptr(t) (already
// offset by this lane's 16*c16 + 4*g); tiles t >= ntl are absent.  tile0 is
// the global tile index of tile 0 (used by the epilogue only).
// The data loads are load(d) when a loader is given (natural-order input),
// else 16-byte loads of tile_ptr(t).
struct DefaultRowsLoad {};

template <int T, bool EPI, int SIGN, int NT = 16, bool CACHED = false, class TilePtr, class Load = DefaultRowsLoad>
__device__ __forceinline__ void rows_body(TilePtr tile_ptr, u32 ntl, u32 tile0, const float2* __restrict__ w4096,
                                          const EpiParams& epi, Load load = {}) {
    static_assert(NT % T == 0 && NT <= 16 && (T < 16 || NT == 16), "tile count per call");
    constexpr int sign = SIGN;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const float sg = sign < 0 ? 1.f : -1.f;

    // table loads first, then the data (see Dit4096Tw); for 1 < T < 16 also
    // the radix-T twiddle base W_{256T}^c of this lane's four positions
    // c = 16*c16 + 4*g + i (its powers W_{256T}^{c*d2} follow by recurrence)
    Dit4096Tw tw;
    dit4096_tw_load(tw, lane, w4096);
    float2 t_base[4];
    if constexpr (T > 1 && T < 16) {
#pragma unroll
        for (int i = 0; i < 4; ++i) t_base[i] = w4096[((16u * c16 + 4u * g + i) * (16u / T)) & 4095u];
    }
    __builtin_amdgcn_sched_barrier(0);
    u32 d[NT][4];
    if constexpr (std::is_same_v<Load, DefaultRowsLoad>) {
        // Unconditional loads in tile order, so that they complete in that
        // order and each tile's first stages wait for that tile only.  Absent
        // tiles (t >= ntl, always whole lines: ntl is a multiple of T) load a
        // copy of the last present tile; the lines they form are computed
        // but never stored, and no line mixes present and absent tiles.
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            load4_nt(d[t], tile_ptr((u32)t < ntl ? t : (int)ntl - 1));
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        load(d);
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
        cf base[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) base[i] = cf{t_base[i].x, sg * t_base[i].y};
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
                cf v[T];
                v[0] = acc[0][i];
                cf p = base[i];  // W_{256T}^{c*tt}
#pragma unroll
                for (int tt = 1; tt < T; ++tt) {
                    v[tt] = cmul(acc[tt][i], p);
                    if (tt + 1 < T) p = cmul(p, base[i]);
                }
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
        // non-temporal, or cached when the next pass reads this output back
        // from the MALL
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if ((u32)t < ntl) {
                if constexpr (CACHED)
                    store4(tile_ptr(t), e[t]);
                else
                    store4_nt(tile_ptr(t), e[t]);
            }
    }
}

template <int T, bool EPI, int SIGN, bool CACHED = false, int NT = 16>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_rows(u32* __restrict__ buf, u32 n_tiles,
                                              const float2* __restrict__ w4096, EpiParams epi,
                                              StartSched sched) {
    const int lane = threadIdx.x & 63;
    // wave index made scalar so that tile addresses are SGPR bases
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 tile0 = wave * (u32)NT;
    if (tile0 >= n_tiles) return;
    start_delay(sched);
    u32* const base = buf + (size_t)tile0 * 256 + 16 * (lane & 15) + 4 * (lane >> 4);
    rows_body<T, EPI, SIGN, NT, CACHED>([&](int t) { return base + t * 256; }, min((u32)NT, n_tiles - tile0), tile0,
                                        w4096, epi);
}

// ----------------------------------------------------------------------------
// Stages 0 and 1 of dit4096 on tiles d2 = TPW*H .. TPW*H + TPW-1 (d[t] =
// tile TPW*H + t), with its stage-1 recurrence, then the LDS rows
// [d2][lds_sigma(c)].  A later part's recurrence is first advanced by the
// updates dit4096 makes after tiles 1 .. TPW*H - 1, so that every product,
// in order, is dit4096's.
template <int SIGN, int TPW, int H>
__device__ __forceinline__ void rows_split_stages(u32 (&d)[TPW][4], const Dit4096Tw& tw, const Dft16Op& op,
                                                  u32* xl) {
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    cf tw1[4], step[4], rec[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tw1[i] = cf{tw.tw1[i].x, sg * tw.tw1[i].y};
        step[i] = cf{tw.step[i].x, sg * tw.step[i].y};
        rec[i] = step[i];
    }
#pragma unroll
    for (int k = 1; k < TPW * H; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) rec[i] = cmul(rec[i], step[i]);
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        constexpr int d2_0 = TPW * H;
        f32x4 re, im;
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(cmul(cf{re[i], im[i]}, tw1[i]));
        dft16_a(d[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf v = {re[i], im[i]};
            if (d2_0 + t > 0) {
                v = cmul(v, rec[i]);
                if (d2_0 + t < 15) rec[i] = cmul(rec[i], step[i]);
            }
            d[t][i] = pack(v);
        }
        // lane (k1 = c16, k0 = 4g + i): columns 4g + 16*c16 + i
        store4(xl + (u32)(d2_0 + t) * 256u + lds_sigma(16u * c16 + 4u * g), d[t]);
    }
}

// Part h (wave-uniform) of a line split over WPL waves: a uniform branch to
// the compile-time instance of rows_split_stages.
template <int SIGN, int WPL>
__device__ __forceinline__ void rows_split_part(u32 h, u32 (&d)[16 / WPL][4], const Dit4096Tw& tw, const Dft16Op& op,
                                                u32* xl) {
    static_assert(WPL == 2 || WPL == 4, "two or four waves per line");
    constexpr int TPW = 16 / WPL;
    if constexpr (WPL == 2) {
        if (h)
            rows_split_stages<SIGN, TPW, 1>(d, tw, op, xl);
        else
            rows_split_stages<SIGN, TPW, 0>(d, tw, op, xl);
    } else {
        switch (h) {
        case 0: rows_split_stages<SIGN, TPW, 0>(d, tw, op, xl); break;
        case 1: rows_split_stages<SIGN, TPW, 1>(d, tw, op, xl); break;
        case 2: rows_split_stages<SIGN, TPW, 2>(d, tw, op, xl); break;
        default: rows_split_stages<SIGN, TPW, 3>(d, tw, op, xl); break;
        }
    }
}

// After the exchange barrier: part h reads columns c = j + 4*c16 + 64q +
// (256/WPL) h (q < 4/WPL) of the 16 points d2 = 4g + i from the exchange
// rows, runs the B-form radix 16 over d2 and, when active, stores X[c +
// 256*k2] (k2 = 4g + i) at line as 16-byte runs of four columns.
template <bool CACHED, int WPL>
__device__ __forceinline__ void rows_split_finish(const u32* xl, u32 h, const Dft16Op& op, u32* __restrict__ line,
                                                  bool active) {
    constexpr int QN = 4 / WPL;
    const int lane = lane_id();
    const u32 c16 = (u32)lane & 15u, g = (u32)lane >> 4;
    const u32 cb = (256u / WPL) * h;
    u32 v[QN][4][4];  // [q][i: point d2 = 4g + i][j: column]
#pragma unroll
    for (int q = 0; q < QN; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) load4(v[q][i], xl + (4u * g + i) * 256u + lds_sigma(4u * c16 + 64u * q + cb));
#pragma unroll
    for (int q = 0; q < QN; ++q) {
        u32 o[4][4];  // [j][i: k2 = 4g + i]
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            u32 a[4] = {v[q][0][j], v[q][1][j], v[q][2][j], v[q][3][j]};
            f32x4 re, im;
            dft16_b(a, op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) o[j][i] = pack(cf{re[i], im[i]});
        }
        if (active) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                u32 s[4] = {o[0][i], o[1][i], o[2][i], o[3][i]};
                u32* const p = line + 256u * (4u * g + i) + 4u * c16 + 64u * q + cb;
                if constexpr (CACHED)
                    store4(p, s);
                else
                    store4_nt(p, s);
            }
        }
    }
}

// k_rows_half<SIGN, CACHED, WPL>: 4096-point lines (no epilogue) split over
// WPL = 2 or 4 waves, 16/WPL tiles each, so that a small batch puts WPL times
// the waves, each with 1/WPL of the load latency, on the chip.  Part h of a
// line loads tiles d2 = (16/WPL) h .. in k_rows<16>'s layout and runs
// stages 0 and 1 of dit4096 on them (rows_split_stages: bit-identical
// results).  The stage-2 inputs go through LDS, [d2][lds_sigma(c)] per line
// (16 KiB): written as 16-byte rows of four columns c = 4g + i + 16*c16,
// read back by part h as columns c = j + 4*c16 + 64q + (256/WPL) h of the
// 16 points d2 = 4g + i, the operand of the B-form radix 16 over d2; output
// X[c + 256*k2], k2 = 4g + i, leaves as 16-byte stores of four columns.
// blockDim.x / (64 WPL) lines per workgroup; a line beyond n_lines loads the
// last line, keeps the barrier and stores nothing.
template <int SIGN, bool CACHED, int WPL>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_rows_half(u32* __restrict__ buf, u32 n_lines,
                                                                   const float2* __restrict__ w4096,
                                                                   StartSched sched) {
    constexpr int TPW = 16 / WPL;
    extern __shared__ u32 lds[];
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 wv = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 h = wv % WPL, ln = wv / WPL;
    const u32 line = blockIdx.x * (blockDim.x / (64u * WPL)) + ln;
    const bool active = line < n_lines;
    start_delay(sched);
    u32* const lbase = buf + (size_t)(active ? line : n_lines - 1u) * 4096u;
    u32* const xl = lds + ln * 4096u;

    Dit4096Tw tw;
    dit4096_tw_load(tw, lane, w4096);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[TPW][4];
    u32* const tin = lbase + (4096u / WPL) * h + 16u * c16 + 4u * g;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        load4_nt(d[t], tin + 256 * t);
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_sched_barrier(0);
    const Dft16Op op = make_dft16(lane, SIGN);
    rows_split_part<SIGN, WPL>(h, d, tw, op, xl);
    __syncthreads();
    rows_split_finish<CACHED, WPL>(xl, h, op, lbase, active);
}

// ----------------------------------------------------------------------------
// k_rows_nat<T>: the first pass of k_rows<T> on natural-order input (the
// digit reversal of the caller's gen_rev done in the load), in -> out, in
// place allowed.  Each wave loads its 16 tiles (whole lines) with 16-byte
// loads, stores them to its own 16 KiB of LDS and reads them back in the
// digit-reversed layout of k_rows: register (t, lane (c16, g), word j) holds
// position p = 256t + 16c16 + 4g + j of the chunk, i.e. natural index
//   m = (t/T)*256T + (t mod T) + T*c16 + 16T*(4g + j)
// of its line.  LDS word of m: m ^ ((m >> 6) & 3) * 0x14 (bits 6-7 into
// bits 2-3 and 4-5): the 16-byte writes stay whole and conflict-free, the
// b32 reads are conflict-free for T = 1 and at most 4-way for T >= 4.  Only
// the wave itself touches its LDS, so no barrier is needed (a wave's LDS
// operations complete in order).  2D (nx_log2 > 0): line l = img*nx + x of
// the input is stored at line img*nx + pos(x), pos = the digit reversal over
// the x radices in reverse order (the inverse of the caller's x reversal),
// out of place only.
__device__ __forceinline__ u32 nat_lds_word(u32 m) { return m ^ (((m >> 6) & 3u) * 0x14u); }

__device__ __forceinline__ u32 rev_pos_pow2(u32 x, const RevSpec& s) {
    u32 digits[12];
    int nd = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i)
        if (i < s.n_digits) {
            const u32 r = (u32)s.radix[i];
            digits[i] = x & (r - 1u);
            x >>= __builtin_ctz(r);
            ++nd;
        }
    u32 idx = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i)
        if (i < s.n_digits) idx = idx * (u32)s.radix[i] + digits[i];
    (void)nd;
    return idx;
}

// Reads a wave's natural-order chunk (ntl tiles, whole lines of 256*T) from
// its LDS slice into the digit-reversed register layout of k_rows.
// The word of (t, j), nat_lds_word(m) with m = (t/T) L + (t%T) + T c16 +
// 16T (4g + j), is (t/T) L + ((t%T) op y_j) with a per-lane y_j (the swizzle
// depends on bits 6-7 of m, which t does not reach): op is + for T <= 4 (y_j
// has no bits below 2) and ^ for T = 8, 16, so the per-element address is a
// constant offset or one XOR.
template <int T, int NT = 16>
__device__ __forceinline__ void nat_rows_read(u32 (&d)[NT][4], const u32* wl, u32 ntl, int lane) {
    constexpr u32 L = 256u * T;
    const u32 c16 = (u32)lane & 15u, g = (u32)lane >> 4;
    u32 y[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const u32 lp = T * c16 + 16u * T * (4u * g + (u32)j);
        y[j] = lp ^ (((lp >> 6) & 3u) * 0x14u);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32 off = T <= 4 ? (u32)(t % T) + y[j] : (y[j] ^ (u32)(t % T));
            d[t][j] = (u32)t < ntl ? wl[(u32)(t / T) * L + off] : 0u;
        }
}

// 16-byte loads of ntl tiles at src (this wave's chunk) into the LDS slice,
// then nat_rows_read.
template <int T>
__device__ __forceinline__ void nat_rows_load(u32 (&d)[16][4], const u32* __restrict__ src, u32* wl, u32 ntl,
                                              int lane) {
    src += 4u * lane;
    u32 v[16][4];
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if ((u32)t < ntl) load4_nt(v[t], src + 256u * t);
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if ((u32)t < ntl) store4(wl + nat_lds_word(256u * t + 4u * lane), v[t]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    nat_rows_read<T>(d, wl, ntl, lane);
}

template <int T, int SIGN, bool CACHED = false>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_rows_nat(const u32* in, u32* out,
                                                                  u32 n_tiles, const float2* __restrict__ w4096,
                                                                  u32 nx_log2, const u32* __restrict__ xpos,
                                                                  StartSched sched) {
    extern __shared__ u32 lds[];
    constexpr u32 L = 256u * T;
    const int lane = threadIdx.x & 63;
    const u32 wib = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + wib;
    const u32 tile0 = wave * 16u;
    if (tile0 >= n_tiles) return;
    start_delay(sched);
    const u32 ntl = min(16u, n_tiles - tile0);
    u32* const wl = lds + wib * 4096u;
    const int c16 = lane & 15, g = lane >> 4;
    auto load = [&](u32 (&d)[16][4]) { nat_rows_load<T>(d, in + (size_t)tile0 * 256u, wl, ntl, lane); };
    const u32 lane_off = 16u * c16 + 4u * g;
    // destination line of each of the wave's 16/T lines, once per line and
    // wave-uniform (2D: the x position from the plan's table)
    constexpr u32 LPW = 16u / T;
    u32 dst_line[LPW];
#pragma unroll
    for (u32 l = 0; l < LPW; ++l) {
        const u32 line = tile0 / T + l;
        u32 dst = line;
        if (nx_log2) {
            const u32 mask = (1u << nx_log2) - 1u;
            dst = (line & ~mask) | xpos[line & mask];
        }
        dst_line[l] = (u32)__builtin_amdgcn_readfirstlane((int)dst);
    }
    auto tile_ptr = [&](int t) -> u32* { return out + (size_t)dst_line[t / T] * L + 256u * ((u32)t % T) + lane_off; };
    rows_body<T, false, SIGN, 16, CACHED>(tile_ptr, ntl, tile0, w4096, EpiParams{}, load);
}

// k_rows_half on natural-order input (a 4096-point row pass of
// tcfft_exec_natural, in -> out, in place allowed): the line's WPL waves
// load its parts with 16-byte loads into 16 KiB of LDS (word nat_lds_word,
// as k_rows_nat), and after a barrier part h reads tiles d2 = (16/WPL) h ..
// of the digit-reversed layout (word y_j ^ d2, see nat_rows_read); the rest
// is k_rows_half's, with the exchange rows in a second 16 KiB.  Every load
// of a line precedes the barrier before its first store, so in place is
// safe.  2D (nx_log2 > 0): input line img*nx + x goes to line img*nx +
// xpos[x].  One line per workgroup.
template <int SIGN, bool CACHED, int WPL>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_rows_half_nat(const u32* in, u32* out, u32 n_lines,
                                                                       const float2* __restrict__ w4096,
                                                                       u32 nx_log2, const u32* __restrict__ xpos,
                                                                       StartSched sched) {
    constexpr int TPW = 16 / WPL;
    extern __shared__ u32 lds[];
    const int lane = lane_id();
    const u32 c16 = (u32)lane & 15u, g = (u32)lane >> 4;
    const u32 h = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 line = blockIdx.x;
    const bool active = line < n_lines;
    start_delay(sched);
    const u32 src_line = active ? line : n_lines - 1u;
    u32 dst_line = src_line;
    if (nx_log2) {
        const u32 mask = (1u << nx_log2) - 1u;
        dst_line = (src_line & ~mask) | xpos[src_line & mask];
    }
    dst_line = (u32)__builtin_amdgcn_readfirstlane((int)dst_line);
    u32* const xin = lds;          // the natural-order line
    u32* const xl = lds + 4096u;   // the exchange rows

    Dit4096Tw tw;
    dit4096_tw_load(tw, lane, w4096);
    __builtin_amdgcn_sched_barrier(0);
    {
        const u32 part = (4096u / WPL) * h;
        const u32* const src = in + (size_t)src_line * 4096u + part + 4u * (u32)lane;
        u32 v[TPW][4];
#pragma unroll
        for (int t = 0; t < TPW; ++t) load4_nt(v[t], src + 256u * t);
#pragma unroll
        for (int t = 0; t < TPW; ++t) store4(xin + nat_lds_word(part + 256u * t + 4u * (u32)lane), v[t]);
    }
    __syncthreads();
    u32 d[TPW][4];
    {
        u32 y[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32 lp = 16u * c16 + 256u * (4u * g + (u32)j);
            y[j] = lp ^ (((lp >> 6) & 3u) * 0x14u);
        }
#pragma unroll
        for (int t = 0; t < TPW; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) d[t][j] = xin[y[j] ^ ((u32)TPW * h + (u32)t)];
    }
    const Dft16Op op = make_dft16(lane, SIGN);
    rows_split_part<SIGN, WPL>(h, d, tw, op, xl);
    __syncthreads();
    rows_split_finish<CACHED, WPL>(xl, h, op, out + (size_t)dst_line * 4096u, active);
}

// ----------------------------------------------------------------------------
// Real-input transforms of N = 512*T points (T = 1..16) on the row kernel.  A
// real signal x[0..N) is read as L = N/2 complex points z[m] = x[2m] + i x[2m+1]
// (the same bytes, 16-byte aligned).  Bins k = 0..L of X = DFT_N(x) follow
// from Z = DFT_L(z) through the even/odd split X = E + W_N^k O:
//   X[k] = ((Z[k] + conj Z[L-k]) - i W_N^k (Z[k] - conj Z[L-k])) / 2,  Z[L] = Z[0].
// One lane handles the pair (k, L-k), k = 0..L/2: with S = Z[k] + conj Z[L-k],
// D = Z[k] - conj Z[L-k] and P = (i/2) W_N^k D,
//   X[k] = S/2 - P,   X[L-k] = conj(S/2 + P)
// (W_N^{L-k} = -conj W_N^k), so every loaded word is used once and one complex
// product serves two bins.  The inverse takes bins 0..L of a Hermitian
// spectrum (imaginary parts of bins 0 and L ignored) and forms
//   Z'[k] = (X[k] + conj X[L-k]) + i W_N^-k (X[k] - conj X[L-k]),  k < L,
// pairwise as Z'[k] = S + P, Z'[L-k] = conj(S - P) with P = i W_N^-k D;
// the unnormalised inverse DFT_L of Z' has re/im = y[2m], y[2m+1] of
// y[t] = sum_{k<N} X[k] e^{2 pi i k t/N} (N times numpy's irfft).
// Complex side: L+1 FP16 complex words per signal, signal s at word s*(L+1).
// W_N^k for k = 64j + lane is wn[lane] * wn[64]^j (wn = W_N^0..64, FP64-made
// FP32 table; the recurrence runs at most 32 steps).

// k_r2c<T>: the k_rows_nat pass with its output kept in the wave's LDS slice
// (natural order, same swizzle), then the pairwise untangle in FP32.  The
// bins go out with ordinary (cached) 4-byte stores: rows start at any 4-byte
// offset, so each store instruction covers parts of three 128-byte lines,
// which L2 merges into whole lines before writing back (non-temporal stores
// of the same pattern: 16384 x 8192, 119.5 us instead of 110.8 us).
template <int T>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_r2c(const u32* __restrict__ in, u32* __restrict__ out,
                                                             u32 n_tiles, const float2* __restrict__ w4096,
                                                             const float2* __restrict__ wn) {
    extern __shared__ u32 lds[];
    constexpr u32 L = 256u * T;
    const int lane = threadIdx.x & 63;
    const u32 wib = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + wib;
    const u32 tile0 = wave * 16u;
    if (tile0 >= n_tiles) return;
    const u32 ntl = min(16u, n_tiles - tile0);
    u32* const wl = lds + wib * 4096u;
    const int c16 = lane & 15, g = lane >> 4;
    const float2 wb = wn[lane], ws = wn[64];
    auto load = [&](u32 (&d)[16][4]) { nat_rows_load<T>(d, in + (size_t)tile0 * 256u, wl, ntl, lane); };
    const u32 lane_off = 16u * c16 + 4u * g;
    auto tile_ptr = [&](int t) -> u32* { return wl + nat_lds_word(256u * (u32)t + lane_off); };
    rows_body<T, false, -1>(tile_ptr, ntl, tile0, w4096, EpiParams{}, load);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const cf step = {ws.x, ws.y};
    const u32 line0 = tile0 / T;
#pragma unroll
    for (u32 i = 0; i < 16u / T; ++i) {
        if (i * T >= ntl) break;
        u32* const dst = out + (size_t)(line0 + i) * (L + 1u);
        cf w = {0.5f * wb.x, 0.5f * wb.y};  // W_N^k / 2
#pragma unroll
        for (u32 j = 0; j < L / 128u; ++j) {
            const u32 k = 64u * j + (u32)lane;  // < L/2
            const cf a = unpack(wl[nat_lds_word(i * L + k)]);
            const cf b = unpack(wl[nat_lds_word(i * L + ((L - k) & (L - 1u)))]);
            const cf hs = {0.5f * (a.x + b.x), 0.5f * (a.y - b.y)};  // S/2
            const cf dif = {a.x - b.x, a.y + b.y};                   // D
            const cf q = cmul(w, dif);                               // W D / 2;  P = i q
            dst[k] = pack(hs.x + q.y, hs.y - q.x);              // S/2 - P
            dst[L - k] = pack(hs.x - q.y, -(hs.y + q.x));       // conj(S/2 + P)
            w = cmul(w, step);
        }
        if (lane == 0) {
            // k = L/2: W_N^{L/2} = -i, so X[L/2] = conj Z[L/2]
            const cf a = unpack(wl[nat_lds_word(i * L + L / 2u)]);
            dst[L / 2u] = pack(a.x, -a.y);
        }
    }
}

// k_c2r<T>: Z' formed pairwise from bins read once from global memory (k
// ascending and L-k descending across lanes), written to the LDS slice in
// natural order, then the k_rows_nat inverse pass to the real output.
template <int T>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_c2r(const u32* __restrict__ in, u32* __restrict__ out,
                                                             u32 n_tiles, const float2* __restrict__ w4096,
                                                             const float2* __restrict__ wn) {
    extern __shared__ u32 lds[];
    constexpr u32 L = 256u * T;
    const int lane = threadIdx.x & 63;
    const u32 wib = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + wib;
    const u32 tile0 = wave * 16u;
    if (tile0 >= n_tiles) return;
    const u32 ntl = min(16u, n_tiles - tile0);
    u32* const wl = lds + wib * 4096u;
    const int c16 = lane & 15, g = lane >> 4;
    const float2 wb = wn[lane], ws = wn[64];
    const u32 line0 = tile0 / T;
    auto load = [&](u32 (&d)[16][4]) {
        const cf step = {ws.x, -ws.y};
#pragma unroll
        for (u32 i = 0; i < 16u / T; ++i) {
            if (i * T >= ntl) break;
            const u32* const src = in + (size_t)(line0 + i) * (L + 1u);
            u32 va[L / 128u], vb[L / 128u];
#pragma unroll
            for (u32 j = 0; j < L / 128u; ++j) {
                va[j] = __builtin_nontemporal_load(src + 64u * j + (u32)lane);
                vb[j] = __builtin_nontemporal_load(src + (L - 64u * j - (u32)lane));
            }
            const u32 vm = lane == 0 ? src[L / 2u] : 0u;
            cf w = {wb.x, -wb.y};  // W_N^-k
#pragma unroll
            for (u32 j = 0; j < L / 128u; ++j) {
                const u32 k = 64u * j + (u32)lane;  // < L/2
                cf a = unpack(va[j]), b = unpack(vb[j]);
                if (j == 0 && lane == 0) a.y = b.y = 0.f;  // bins 0 and L: real parts only
                const cf sum = {a.x + b.x, a.y - b.y};   // S
                const cf dif = {a.x - b.x, a.y + b.y};   // D
                const cf q = cmul(w, dif);               // P = i q
                wl[nat_lds_word(i * L + k)] = pack(sum.x - q.y, sum.y + q.x);                  // S + P
                if (k != 0) wl[nat_lds_word(i * L + (L - k))] = pack(sum.x + q.y, q.x - sum.y);  // conj(S - P)
                w = cmul(w, step);
            }
            if (lane == 0) {
                // k = L/2: W_N^{-L/2} = i, so Z'[L/2] = 2 conj X[L/2]
                const cf a = unpack(vm);
                wl[nat_lds_word(i * L + L / 2u)] = pack(2.f * a.x, -2.f * a.y);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        nat_rows_read<T>(d, wl, ntl, lane);
    };
    const u32 lane_off = 16u * c16 + 4u * g;
    auto tile_ptr = [&](int t) -> u32* { return out + (size_t)(tile0 + (u32)t) * 256u + lane_off; };
    rows_body<T, false, +1>(tile_ptr, ntl, tile0, w4096, EpiParams{}, load);
}

// ----------------------------------------------------------------------------
// k_cols<T>: strided pass of radix R = 16*T at element step s over a buffer
// viewed as [outer][R][s].  A wave owns 16 adjacent columns of one block.
// Rows j = j0 + 16*j1 are digit-reversed for [16, T]: stage 0 is a DFT over
// j0 (adjacent rows), stage 1 a radix-T over j1 with W_{16T}^{k0*j1}.
// Lane (c16, g) word i of tile j1 holds row j0 = 4g+i, column c16.
template <int T, bool EPI>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_cols(u32* __restrict__ buf, u32 s_log2, u32 n_waves, int sign,
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
    constexpr u32 R = 16 * T;
    const size_t blk_base = (size_t)blk * R * s + col0 + c16;

    // Table values before the data (global loads complete in issue order):
    //   step[i] = W_{16T}^{k0}, k0 = 4g + i: the radix-T (T < 16) or
    //   stage-1 (T = 16) twiddle W_{16T}^{k0*t} is its t-th power;
    //   epilogue: output row k0 + 16*t (T < 16), t + 16*k0 (T = 16) or k0
    //   (T = 1) gets F0[i] * S^t with F0[i] = W_M^{(xc + k0*sa*u)*rj}
    //   (u = 16 for T = 16, else 1) and S = W_M^{sa*rj*(T == 16 ? 1 : 16)}.
    float2 t_step[4];
    if constexpr (T > 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) t_step[i] = w4096[(4 * g + i) * (256 / T)];
    }
    struct RawW {
        float2 h, l;
    };
    RawW r_f0[4], r_s;
    if constexpr (EPI) {
        const u32 rj = epi_rev(epi, blk);
        const u32 sa = s >> epi.shift, xc = (col0 + (u32)c16) >> epi.shift;
        const u32 u = T == 16 ? 16u : 1u;
        auto ld = [&](u32 m) -> RawW {
            m &= epi.m_mask;
            return RawW{epi.hi[m >> epi.lo_bits], epi.lo[m & ((1u << epi.lo_bits) - 1u)]};
        };
#pragma unroll
        for (int i = 0; i < 4; ++i) r_f0[i] = ld((xc + (4u * g + i) * u * sa) * rj);
        r_s = ld(sa * rj * (T == 16 ? 1u : 16u));
    }
    __builtin_amdgcn_sched_barrier(0);
    u32 d[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = buf[blk_base + (size_t)(16 * t + 4 * g + i) * s];
    __builtin_amdgcn_sched_barrier(0);

    const Dft16Op op = make_dft16(lane, sign);
    cf step[4];
    if constexpr (T > 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) step[i] = cf{t_step[i].x, sg * t_step[i].y};
    }
    auto rawv = [&](const RawW& r) -> cf { return cmul(cf{r.h.x, sg * r.h.y}, cf{r.l.x, sg * r.l.y}); };
    cf F0[4], S = {1.f, 0.f};
    if constexpr (EPI) {
#pragma unroll
        for (int i = 0; i < 4; ++i) F0[i] = rawv(r_f0[i]);
        if constexpr (T > 1) S = rawv(r_s);
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
            cf p = step[i];  // W_{16T}^{k0*t}
#pragma unroll
            for (int t = 1; t < T; ++t) {
                v[t] = cmul(acc[t][i], p);
                if (t + 1 < T) p = cmul(p, step[i]);
            }
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
        // T == 16: stage 0 -> (g,i) = k0, tile = j1; twiddle W_256^{k0*j1}
        // (recurrence over the tiles); transpose to (g,i) = j1, tile = k0;
        // stage 1 -> (g,i) = k1.
        cf rec[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) rec[i] = step[i];
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {re[i], im[i]};
                if (t) {
                    v = cmul(v, rec[i]);
                    if (t < 15) rec[i] = cmul(rec[i], step[i]);
                }
                d[t][i] = pack(v);
            }
        }
        u32 e[16][4];
        transpose_gi_tile(d, e);
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
// k_cols_w<T>: the strided radix-16T pass of k_cols<T> (T = 8, 16; no
// epilogue) for a workgroup of 4 waves that owns 64 adjacent columns of one
// [16T][s] block (s >= 64), with 16-byte global accesses: the workgroup loads
// the 16T rows x 256 bytes into LDS (16-byte lanes, four rows per wave
// instruction), each wave reads its 16 columns in the B-form layout of
// k_cols<T> (lane (c16, g) word i of tile j1 = row j0 + 16*j1, j0 = 4g+i),
// computes the same stages, writes its outputs back into the tile and the
// workgroup stores the rows with 16-byte lanes.  LDS word of (row r, column c):
// 64r + (c ^ 16*(((r >> 2) ^ (r >> 6)) & 3)): the 4-byte column reads (rows
// 4g apart) and writes (rows 64g apart) spread the four lane groups over
// different 16-bank groups, and a 16-byte piece of 4 columns stays
// contiguous.
__device__ __forceinline__ u32 colsw_word(u32 r, u32 c) { return 64u * r + (c ^ ((((r >> 2) ^ (r >> 6)) & 3u) << 4)); }

template <int T, int SIGN>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_cols_w(u32* __restrict__ buf, u32 s_log2, u32 n_groups,
                                                                const float2* __restrict__ w4096) {
    static_assert(T == 8 || T == 16, "k_cols_w: radix 128 or 256");
    constexpr u32 R = 16u * T;
    extern __shared__ u32 lds[];
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const u32 grp = blockIdx.x;
    if (grp >= n_groups) return;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 tid = threadIdx.x;
    const u32 w = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 cg_log2 = s_log2 - 6u;  // 64-column groups per block
    const u32 blk = grp >> cg_log2;
    const u32 col0 = (grp & ((1u << cg_log2) - 1u)) * 64u;
    u32* const base = buf + (((size_t)blk * R) << s_log2) + col0;
    // load phase: thread (rr = tid >> 4, cq = tid & 15) takes rows rr + 16k,
    // columns 4cq..4cq+3
    const u32 rr = tid >> 4, cq = tid & 15u;
    float2 t_step[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t_step[i] = w4096[(256u / T) * (4u * g + i)];  // W_R^{k0}
    __builtin_amdgcn_sched_barrier(0);
    {
        u32 v[T][4];
#pragma unroll
        for (int k = 0; k < T; ++k) load4_nt(v[k], base + ((size_t)(rr + 16u * k) << s_log2) + 4u * cq);
#pragma unroll
        for (int k = 0; k < T; ++k) store4(lds + colsw_word(rr + 16u * k, 4u * cq), v[k]);
    }
    __syncthreads();
    u32 d[T][4];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = lds[colsw_word(16u * t + 4u * g + i, 16u * w + c16)];
    const Dft16Op op = make_dft16(lane, SIGN);
    cf step[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) step[i] = cf{t_step[i].x, sg * t_step[i].y};
    if constexpr (T == 16) {
        {
            cf rec[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) rec[i] = step[i];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                f32x4 re, im;
                dft16_b(d[t], op, re, im);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    cf v = {re[i], im[i]};
                    if (t) {
                        v = cmul(v, rec[i]);
                        if (t < 15) rec[i] = cmul(rec[i], step[i]);
                    }
                    d[t][i] = pack(v);
                }
            }
        }
        u32 e[16][4];
        transpose_gi_tile(d, e);  // (g,i) = j1, tile = k0
        __syncthreads();          // every wave has read its columns
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(e[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                lds[colsw_word((u32)t + 16u * (4u * g + i), 16u * w + c16)] = pack(cf{re[i], im[i]});
        }
    } else {
        // stage 0 per tile (rows j0 = 4g + i, tile j1), then radix T across
        // the tiles with W_R^{k0*j1}: output row k0 + 16*k1 (k0 = 4g + i)
        cf acc[T][4];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[t][i] = cf{re[i], im[i]};
        }
        __syncthreads();  // every wave has read its columns
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf v[T];
            v[0] = acc[0][i];
            cf p = step[i];
#pragma unroll
            for (int t = 1; t < T; ++t) {
                v[t] = cmul(acc[t][i], p);
                if (t + 1 < T) p = cmul(p, step[i]);
            }
            small_dft<T>(v, SIGN);
#pragma unroll
            for (int t = 0; t < T; ++t) lds[colsw_word(4u * g + i + 16u * t, 16u * w + c16)] = pack(v[t]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < T; ++k) {
        u32 v[4];
        load4(v, lds + colsw_word(rr + 16u * k, 4u * cq));
        store4_nt(base + ((size_t)(rr + 16u * k) << s_log2) + 4u * cq, v);
    }
}

// ----------------------------------------------------------------------------
// k_cols_q<T>: the strided radix-16T pass (T = 1, 2, 4) with four adjacent
// columns per lane (16-byte loads and stores, 64 columns per wave, 4x the
// bytes in flight of k_cols), for passes without an epilogue twiddle.
// Lane (c16, g) owns columns col0 + 4*c16 + q, q = 0..3; MFMA q takes
// column 4*c16 + q as its column c16.
template <int T>
__global__ __launch_bounds__(256) TCFFT_MFMA_BUDGET void k_cols_q(u32* __restrict__ buf, u32 s_log2, u32 n_waves, int sign,
                                                const float2* __restrict__ w4096) {
    static_assert(T == 1 || T == 2 || T == 4, "k_cols_q: T = 1, 2, 4");
    const int lane = threadIdx.x & 63;
    const int c16 = lane & 15, g = lane >> 4;
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= n_waves) return;
    const u32 s = 1u << s_log2;
    const u32 col_groups = s >> 6;
    const u32 blk = wave / col_groups;
    const u32 col0 = (wave - blk * col_groups) * 64u;
    constexpr u32 R = 16 * T;
    const float sg = sign < 0 ? 1.f : -1.f;
    // W_{16T}^{k0*t}, k0 = 4g + i, t >= 1: loaded before the data
    float2 tw[4][T > 1 ? T - 1 : 1];
    if constexpr (T > 1) {
        constexpr int STRIDE = 256 / T;  // W_{16T}^m = W_4096^{m*STRIDE}
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int t = 1; t < T; ++t) tw[i][t - 1] = w4096[(((4 * g + i) * t) * STRIDE) & 4095];
    }
    __builtin_amdgcn_sched_barrier(0);
    u32* const base = buf + (size_t)blk * R * s + col0 + 4u * c16;
    u32 d[T][4][4];  // [tile][row word i][column q]
    // cached loads: the pass before may have left this data in the MALL
    // (2^20 x 16: 47.6 -> 46.6 us over the exec; no change where it has not)
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) load4(d[t][i], base + (size_t)(16 * t + 4 * g + i) * s);
    __builtin_amdgcn_sched_barrier(0);
    const Dft16Op op = make_dft16(lane, sign);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cf acc[T][4];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const u32 a[4] = {d[t][0][q], d[t][1][q], d[t][2][q], d[t][3][q]};
            f32x4 re, im;
            dft16_b(a, op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[t][i] = cf{re[i], im[i]};
        }
        if constexpr (T > 1) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v[T];
                v[0] = acc[0][i];
#pragma unroll
                for (int t = 1; t < T; ++t) v[t] = cmul(acc[t][i], cf{tw[i][t - 1].x, sg * tw[i][t - 1].y});
                small_dft<T>(v, sign);
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t][i] = v[t];
            }
        }
        // output row k0 + 16t (k0 = 4g + i) of column q
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) d[t][i][q] = pack(acc[t][i]);
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) store4_nt(base + (size_t)(4 * g + i + 16 * t) * s, d[t][i]);
}

// ----------------------------------------------------------------------------
// k_cols_big<W>: the strided pass of radix R = 256*W (W = 2, 4, 8) at element
// step s in one HBM round trip, as the last pass of its axis (no epilogue).
// A workgroup of W waves owns 16 adjacent columns of one [R][s] block.  Rows
// j = j0 + 16*j1 + 256*j2 are digit-reversed for [16, 16, W]; wave w loads
// the 256 rows with j2 = w (lane (c16, g) word i of tile j1: row j0 = 4g+i,
// column c16) and computes their 256-point DFT times the radix-W stage's
// twiddle W_{256W}^{m*w}, m = k0 + 16*k1, at no per-element cost:
//   stage 0: DFT-16 over j0 (B-form) -> (g,i) = k0, times W_256^{k0*j1} *
//            W_{256W}^{k0*w} (a recurrence over the tiles j1 that starts at
//            W_{256W}^{k0*w}); transpose to tile = k0, (g,i) = j1;
//   stage 1: DFT-16 over j1 whose operand rows are scaled by
//            W_{256W}^{16*k1*w} -> tile k0, (g,i) = k1.
// Exchange through LDS laid out [source wave][tile][lane][4 words] per round
// (b128 both ways, conflict-free), each wave taking TPW tiles of every source
// per round, then the radix-W stage over the sources (VALU) writes row
// k0 + 16*k1 + 256*k2.  Workgroups are mapped so that those on one XCD
// (block mod 8) take adjacent column groups, whose 64-byte row segments share
// 128-byte lines in that XCD's L2.
template <int W>
struct ColsBigCfg {
    static constexpr int ROUNDS = W >= 4 ? 2 : 1;
    static constexpr int TPR = 16 / ROUNDS;  // tiles exchanged per round
    static constexpr int TPW = TPR / W;      // tiles per reader wave per round
    static constexpr int LDS_WORDS = W * TPR * 256;
};

template <int W, int SIGN>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_cols_big(u32* __restrict__ buf, u32 s_log2,
                                                                     u32 n_groups,
                                                                     const float2* __restrict__ w4096,
                                                                     StartSched sched) {
    using C = ColsBigCfg<W>;
    static_assert(W == 2 || W == 4 || W == 8, "k_cols_big: W = 2, 4, 8");
    extern __shared__ u32 lds[];
    start_delay(sched);
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 w = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    u32 grp = blockIdx.x;
    if ((n_groups & 7u) == 0u) grp = (blockIdx.x & 7u) * (n_groups >> 3) + (blockIdx.x >> 3);
    const u32 cg_log2 = s_log2 - 4u;  // column groups of 16 per block: s / 16
    const u32 blk = grp >> cg_log2;
    const u32 col0 = (grp & ((1u << cg_log2) - 1u)) * 16u;
    // wave-uniform bases and 32-bit lane byte offsets (at most 192 rows of
    // s <= 2^19 elements into the block), so that no load or store needs a
    // 64-bit per-lane address
    u32* const blk_u = buf + (((size_t)blk * (256u * W)) << s_log2) + col0;
    const u32 lane_ld = ((u32)c16 + ((4u * g) << s_log2)) * 4u;
    const u32 lane_st = ((u32)c16 + ((64u * g) << s_log2)) * 4u;

    // table values before the data (global loads complete in issue order)
    float2 t_step[4], t_rec0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const u32 k0 = 4u * g + i;
        t_step[i] = w4096[16u * k0];                    // W_256^{k0}
        t_rec0[i] = w4096[(k0 * w * (16u / W)) & 4095u];  // W_{256W}^{k0*w}
    }
    const float2 t_s1 = w4096[((u32)c16 * w * (256u / W)) & 4095u];  // W_{256W}^{16*c16*w}
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const u32* up = blk_u + ((size_t)(16u * t + i + 256u * w) << s_log2);
            d[t][i] = ld_at<u32>(up, lane_ld);
        }
    __builtin_amdgcn_sched_barrier(0);

    const Dft16Op op = make_dft16(lane, SIGN);
    {
        cf step[4], rec[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            step[i] = cf{t_step[i].x, sg * t_step[i].y};
            rec[i] = cf{t_rec0[i].x, sg * t_rec0[i].y};
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                d[t][i] = pack(cmul(cf{re[i], im[i]}, rec[i]));
                if (t < 15) rec[i] = cmul(rec[i], step[i]);
            }
        }
    }
    u32 e[16][4];
    transpose_gi_tile(d, e);  // tile = k0, (g,i) = j1
    const Dft16Op op1 = make_dft16_scaled(lane_id(), SIGN, cf{t_s1.x, sg * t_s1.y});
#pragma unroll
    for (int r = 0; r < C::ROUNDS; ++r) {
        if (r) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int tt = 0; tt < C::TPR; ++tt) {
            f32x4 re, im;
            dft16_b(e[r * C::TPR + tt], op1, re, im);
            u32 q[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) q[i] = pack(cf{re[i], im[i]});
            store4(lds + ((w * C::TPR + tt) * 64u + (u32)lane) * 4u, q);
        }
        __syncthreads();
#pragma unroll
        for (int ti = 0; ti < C::TPW; ++ti) {
            const u32 tt = w * C::TPW + ti;
            u32 v[W][4];
#pragma unroll
            for (int src = 0; src < W; ++src) load4(v[src], lds + (((u32)src * C::TPR + tt) * 64u + (u32)lane) * 4u);
            const u32 k0 = (u32)r * C::TPR + tt;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf x[W];
#pragma unroll
                for (int src = 0; src < W; ++src) x[src] = unpack(v[src][i]);
                small_dft<W>(x, SIGN);
                // row k0 + 16 (4g + i) + 256 k2: the 64 g part is the lane's
#pragma unroll
                for (int k2 = 0; k2 < W; ++k2) {
                    u32* up = blk_u + ((size_t)(k0 + 16u * i + 256u * k2) << s_log2);
                    st_at<u32>(up, lane_st, pack(x[k2]));
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// k_cols_bigw<W>: k_cols_big<W> for 32 adjacent columns per workgroup of 2W
// waves, with 16-byte global accesses (used for W = 2; at W = 4 the 144 KiB
// tile leaves one workgroup per CU, and it measured slower than k_cols_big:
// 2D 1024^2 x 8 32.3 -> 34.5 us, 2^26 x 4 972 -> 1098 us).  The workgroup loads
// the block's 256W rows x 128 bytes into an LDS tile (16-byte lanes, 16W rows
// per round); wave (w, q) = (row group, 16-column group) reads the 256 rows
// j2 = w of its 16 columns in k_cols_big's B-form layout and runs the same two
// stages; after a barrier the tile region holds the radix-W exchange
// ([q][source wave][tile][lane][4 words], one round), each wave's radix-W
// outputs stay in registers until a second barrier, go back into the tile and
// leave with 16-byte non-temporal stores of whole 128-byte row pieces.  Tile
// word of (row r, column c): 36r + (c ^ 16*((r >> 6) & 1)) (rows of 144
// bytes): the B-form column reads (rows 4g apart) fall on four different
// 16-bank groups, the output writes (rows 64g apart) on two.
template <int W>
struct ColsBigWCfg {
    static constexpr u32 ROWS = 256u * W;
    static constexpr u32 PITCH = 36u;
    static constexpr u32 TILE_WORDS = ROWS * PITCH;
    static constexpr u32 XCH_WORDS = 2u * W * 16u * 256u;  // [q][src][tile][lane][4]
    static constexpr u32 LDS_WORDS = TILE_WORDS > XCH_WORDS ? TILE_WORDS : XCH_WORDS;
};
__device__ __forceinline__ u32 colsbw_word(u32 r, u32 c) { return 36u * r + (c ^ (((r >> 6) & 1u) << 4)); }

template <int W, int SIGN>
__global__ __launch_bounds__(128 * W) TCFFT_MFMA_BUDGET void k_cols_bigw(u32* __restrict__ buf, u32 s_log2,
                                                                       u32 n_groups,
                                                                       const float2* __restrict__ w4096,
                                                                       StartSched sched) {
    static_assert(W == 2 || W == 4, "k_cols_bigw: W = 2, 4 (the tile of W = 8 exceeds the LDS)");
    using C = ColsBigWCfg<W>;
    extern __shared__ u32 lds[];
    const u32 grp = blockIdx.x;
    if (grp >= n_groups) return;
    start_delay(sched);
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 tid = threadIdx.x;
    const u32 wv = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 w = wv % W, q = wv / W;
    const u32 cg_log2 = s_log2 - 5u;  // 32-column groups per block
    const u32 blk = grp >> cg_log2;
    const u32 col0 = (grp & ((1u << cg_log2) - 1u)) * 32u;
    u32* const base = buf + (((size_t)blk * C::ROWS) << s_log2) + col0;
    // load phase: thread (rr = tid >> 3, cq = tid & 7) takes rows rr + 16W*k,
    // columns 4cq..4cq+3
    const u32 rr = tid >> 3, cq = tid & 7u;

    // table values before the data (global loads complete in issue order)
    float2 t_step[4], t_rec0[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const u32 k0 = 4u * g + i;
        t_step[i] = w4096[16u * k0];                    // W_256^{k0}
        t_rec0[i] = w4096[(k0 * w * (16u / W)) & 4095u];  // W_{256W}^{k0*w}
    }
    const float2 t_s1 = w4096[((u32)c16 * w * (256u / W)) & 4095u];  // W_{256W}^{16*c16*w}
    __builtin_amdgcn_sched_barrier(0);
    {
        u32 v[16][4];
#pragma unroll
        for (int k = 0; k < 16; ++k) load4_nt(v[k], base + ((size_t)(rr + 16u * W * k) << s_log2) + 4u * cq);
#pragma unroll
        for (int k = 0; k < 16; ++k) store4(lds + colsbw_word(rr + 16u * W * k, 4u * cq), v[k]);
    }
    __syncthreads();
    u32 d[16][4];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = lds[colsbw_word(4u * g + i + 16u * t + 256u * w, 16u * q + c16)];

    const Dft16Op op = make_dft16(lane, SIGN);
    {
        cf step[4], rec[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            step[i] = cf{t_step[i].x, sg * t_step[i].y};
            rec[i] = cf{t_rec0[i].x, sg * t_rec0[i].y};
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 re, im;
            dft16_b(d[t], op, re, im);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                d[t][i] = pack(cmul(cf{re[i], im[i]}, rec[i]));
                if (t < 15) rec[i] = cmul(rec[i], step[i]);
            }
        }
    }
    u32 e[16][4];
    transpose_gi_tile(d, e);  // tile = k0, (g,i) = j1
    const Dft16Op op1 = make_dft16_scaled(lane_id(), SIGN, cf{t_s1.x, sg * t_s1.y});
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_b(e[t], op1, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) e[t][i] = pack(cf{re[i], im[i]});
    }
    __syncthreads();  // every wave has read its input columns: the tile region is free
    u32* const xq = lds + q * (W * 16u * 256u);
#pragma unroll
    for (int t = 0; t < 16; ++t) store4(xq + ((w * 16u + (u32)t) * 64u + (u32)lane) * 4u, e[t]);
    __syncthreads();
    // radix-W over the sources for tiles k0 = w*TPW + ti; outputs row
    // k0 + 16 i + 64 g + 256 k2, column 16 q + c16
    constexpr u32 TPW = 16u / W;
    u32 o[TPW][4][W];
#pragma unroll
    for (u32 ti = 0; ti < TPW; ++ti) {
        const u32 k0 = w * TPW + ti;
        u32 v[W][4];
#pragma unroll
        for (int src = 0; src < W; ++src) load4(v[src], xq + (((u32)src * 16u + k0) * 64u + (u32)lane) * 4u);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            cf x[W];
#pragma unroll
            for (int src = 0; src < W; ++src) x[src] = unpack(v[src][i]);
            small_dft<W>(x, SIGN);
#pragma unroll
            for (int k2 = 0; k2 < W; ++k2) o[ti][i][k2] = pack(x[k2]);
        }
    }
    __syncthreads();  // the exchange has been read: the tile region takes the outputs
#pragma unroll
    for (u32 ti = 0; ti < TPW; ++ti)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int k2 = 0; k2 < W; ++k2)
                lds[colsbw_word(w * TPW + ti + 16u * i + 64u * g + 256u * k2, 16u * q + c16)] = o[ti][i][k2];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        u32 v[4];
        load4(v, lds + colsbw_word(rr + 16u * W * k, 4u * cq));
        store4_nt(base + ((size_t)(rr + 16u * W * k) << s_log2) + 4u * cq, v);
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

template <int W, bool EPI, int SIGN, bool CACHED = false>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_big(u32* __restrict__ buf, const float2* __restrict__ w4096,
                                                const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo,
                                                u32 wn_lo_bits, EpiParams epi, StartSched sched) {
    extern __shared__ u32 lds[];
    start_delay(sched);
    big_body<W, BigCfg<W>::ROUNDS, EPI, SIGN, CACHED>(buf + (size_t)blockIdx.x * BigCfg<W>::N, threadIdx.x >> 6,
                                                     lds, blockIdx.x, w4096, wn_hi, wn_lo, wn_lo_bits, epi);
}

// ----------------------------------------------------------------------------
// k_big_nat<W>: k_big<W> (W = 2, 4, 8, 16) on natural-order input, in -> out (in
// place allowed: a workgroup reads its whole signal before its first store).
// Each wave loads its natural 4096-element chunk with 16-byte loads and stores
// it to LDS (the signal, W*16 KiB, fits at once); after a barrier wave w reads
// the chunk k_big loads in the drop-in order: register (t, lane (c16, g),
// word j) = position 4096w + 256t + 16c16 + 4g + j, natural index
// w + W*(t + 16*c16 + 256*(4g + j)).  The LDS word XORs two 2-bit fields of
// the index into bits 2-3 and 4-5 (NatBigSwz): the 16-byte writes stay whole
// and conflict-free, the b32 reads drop from 32..64-way to 4-way conflicts.
// 2D (nx_log2 > 0, out of place): line img*nx + x goes to img*nx + pos(x).
template <int W>
struct NatBigSwz;
template <>
struct NatBigSwz<2> {
    static constexpr u32 A = 7, B = 6;
};
template <>
struct NatBigSwz<4> {
    static constexpr u32 A = 6, B = 8;
};
template <>
struct NatBigSwz<8> {
    static constexpr u32 A = 7, B = 9;
};

template <int W>
__device__ __forceinline__ u32 nat_big_word(u32 m) {
    return m ^ (((m >> NatBigSwz<W>::A) & 3u) << 2) ^ (((m >> NatBigSwz<W>::B) & 3u) << 4);
}

// W = 16 (65536 points, 256 KiB per signal) goes through 128 KiB of LDS in
// two rounds split by index bit 11: writer tile bit 3, reader lane bit
// c16[3].  Round word: bit 11 removed, then bits 8-9 into 2-3 and 13-14 into
// 4-5 (writes conflict-free, reads at most 2-way among the 32 active lanes).
__device__ __forceinline__ u32 nat16_word(u32 n) {
    const u32 m = (n & 2047u) | ((n >> 12) << 11);
    return m ^ (((m >> 8) & 3u) << 2) ^ (((m >> 13) & 3u) << 4);
}

// Word of natural index n in a staging round that holds the indices with
// bit B of n equal to the round number: bit B removed, then bits 8-9 into
// 2-3 and 13-14 into 4-5 (a bijection that keeps 16-byte chunks whole).
template <int B>
__device__ __forceinline__ u32 nat_split_word(u32 n) {
    const u32 m = (n & ((1u << B) - 1u)) | ((n >> (B + 1)) << B);
    return m ^ (((m >> 8) & 3u) << 2) ^ (((m >> 13) & 3u) << 4);
}

// nat_split_word<B>(p + q) for p < 2^LW and a lane index q whose bits below
// LW are zero (LW <= 8 <= B): the swizzle bits come from q alone, so the
// word is base + (p ^ xlo) with base and xlo computed once per lane.
template <int B, int LW>
__device__ __forceinline__ void nat_split_lane(u32 q, u32& base, u32& xlo) {
    const u32 m = (q & ((1u << B) - 1u)) | ((q >> (B + 1)) << B);
    const u32 x = (((m >> 8) & 3u) << 2) ^ (((m >> 13) & 3u) << 4);
    xlo = x & ((1u << LW) - 1u);
    base = m ^ (x & ~((1u << LW) - 1u));
}

template <int W, int SIGN, bool CACHED = false>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_big_nat(
    const u32* in, u32* out, const float2* __restrict__ w4096,
    const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo, u32 wn_lo_bits, u32 nx_log2, RevSpec x_pos,
    StartSched sched) {
    extern __shared__ u32 lds[];
    constexpr u32 N = 4096u * W;
    start_delay(sched);
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 w = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 sig = blockIdx.x;
    BigTw bt;
    u32 d[16][4];
    if constexpr (W < 16) {
        big_tw_load<W, false>(bt, lane, (int)w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, EpiParams{});
        __builtin_amdgcn_sched_barrier(0);
        {
            u32 v[16][4];
            const u32* src = in + (size_t)sig * N + 4096u * w + 4u * lane;
#pragma unroll
            for (int t = 0; t < 16; ++t) load4_nt(v[t], src + 256u * t);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < 16; ++t) store4(lds + nat_big_word<W>(4096u * w + 256u * t + 4u * lane), v[t]);
        }
        __syncthreads();
        // reader word of register (t, j): nat_big_word<W>(w + W (t + 16 c16 +
        // 256 (4g + j))) = lane_base + ((w + W t) ^ xlo) + 256 W j: the swizzle
        // bits come from the lane's part q = 16W c16 + 1024W g alone, its bits
        // below log2(16W) (xlo) meet w + W t, the rest (q's) do not
        constexpr u32 LW = W == 2 ? 5u : W == 4 ? 6u : 7u;  // log2(16W)
        const u32 q = 16u * W * (u32)c16 + 1024u * W * (u32)g;
        const u32 x = (((q >> NatBigSwz<W>::A) & 3u) << 2) ^ (((q >> NatBigSwz<W>::B) & 3u) << 4);
        const u32 xlo = x & ((1u << LW) - 1u), lane_base = q ^ (x & ~((1u << LW) - 1u));
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const u32* a = lds + lane_base + ((w + W * (u32)t) ^ xlo);
#pragma unroll
            for (int j = 0; j < 4; ++j) d[t][j] = a[256u * W * j];
        }
    } else {
        // data first, the twiddle tables after the exchange: the pending
        // half of the loaded data and the reader registers fit the budget
        u32 v[16][4];
        const u32* src = in + (size_t)sig * N + 4096u * w + 4u * lane;
#pragma unroll
        for (int t = 0; t < 16; ++t) load4_nt(v[t], src + 256u * t);
        // reader word of register (t, j): nat16_word(w + 16 (t + 16 c16 +
        // 256 (4g + j))) = lane_base + 16 (t ^ g) + 2048 j, computed once per
        // lane (the per-element form cost about 1600 VALU instructions a wave)
        const u32 lane_base = ((u32)w ^ (4u * ((u32)c16 & 3u))) + 256u * ((u32)c16 & 7u) + 8192u * (u32)g;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r) __syncthreads();  // round 0's reads are done
#pragma unroll
            for (int tt = 0; tt < 8; ++tt)
                store4(lds + nat16_word(4096u * w + 256u * (8u * r + tt) + 4u * lane), v[8 * r + tt]);
            __syncthreads();
            if ((u32)(c16 >> 3) == (u32)r) {
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const u32* a = lds + lane_base + 16u * ((u32)t ^ (u32)g);
#pragma unroll
                    for (int j = 0; j < 4; ++j) d[t][j] = a[2048u * j];
                }
            }
        }
        big_tw_load<W, false>(bt, lane_id(), (int)w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, EpiParams{});
    }
    __syncthreads();  // every wave has read its chunk before the exchange reuses the LDS
    u32 dst = sig;
    if (nx_log2) {
        const u32 x = sig & ((1u << nx_log2) - 1u);
        dst = (sig & ~((1u << nx_log2) - 1u)) | rev_pos_pow2(x, x_pos);
    }
    const Dft16Op op = make_dft16(lane, SIGN);
    u32 e[16][4];
    big_phase1<SIGN>(d, e, bt, op);
    big_phase2<W, BigCfg<W>::ROUNDS, false, SIGN, false, CACHED>(e, (int)w, lds, out + (size_t)dst * N, bt, op,
                                                                 true);
}

// ----------------------------------------------------------------------------
// Real-input transforms above 65536 points: the L = N/2-point complex
// transform runs as its own passes and k_pairs<SIGN> is the even/odd split by
// itself, one lane per bin pair (k, L-k), k < L/2 (the lane with k = 0 also
// takes bin L/2), with W_N^k = hi[k >> lo_bits] * lo[k & mask] from the
// two-level table of W_N (same formulas as k_r2c / k_c2r).
//   SIGN -1: Z (natural, L words per signal) -> bins 0..L (L+1 words per signal)
//   SIGN +1: bins 0..L -> Z' (L words per signal)
template <int SIGN>
__global__ __launch_bounds__(256) void k_pairs(const u32* __restrict__ src, u32* __restrict__ dst, u32 l_log2,
                                               u32 n_pairs, const float2* __restrict__ hi,
                                               const float2* __restrict__ lo, u32 lo_bits) {
    const u32 t = blockIdx.x * 256u + threadIdx.x;
    if (t >= n_pairs) return;
    const u32 L = 1u << l_log2, half = L >> 1;
    const size_t s = t >> (l_log2 - 1u);
    const u32 k = t & (half - 1u);
    const float2 th = hi[k >> lo_bits], tl = lo[k & ((1u << lo_bits) - 1u)];
    const cf w = cmul(cf{th.x, th.y}, cf{tl.x, tl.y});  // W_N^k
    if constexpr (SIGN < 0) {
        const u32* const z = src + s * L;
        u32* const x = dst + s * (L + 1u);
        const cf a = unpack(z[k]);
        const cf b = unpack(z[(L - k) & (L - 1u)]);
        const cf hs = {0.5f * (a.x + b.x), 0.5f * (a.y - b.y)};  // S/2
        const cf dif = {a.x - b.x, a.y + b.y};                   // D
        const cf q = cmul(cf{0.5f * w.x, 0.5f * w.y}, dif);      // P = i q
        x[k] = pack(hs.x + q.y, hs.y - q.x);
        x[L - k] = pack(hs.x - q.y, -(hs.y + q.x));
        if (k == 0) {
            const cf m = unpack(z[half]);
            x[half] = pack(m.x, -m.y);
        }
    } else {
        const u32* const x = src + s * (L + 1u);
        u32* const z = dst + s * L;
        cf a = unpack(x[k]), b = unpack(x[L - k]);
        if (k == 0) a.y = b.y = 0.f;  // bins 0 and L: real parts only
        const cf sum = {a.x + b.x, a.y - b.y};
        const cf dif = {a.x - b.x, a.y + b.y};
        const cf q = cmul(cf{w.x, -w.y}, dif);  // W_N^-k D; P = i q
        z[k] = pack(sum.x - q.y, sum.y + q.x);
        if (k != 0) z[L - k] = pack(sum.x + q.y, q.x - sum.y);
        if (k == 0) {
            const cf m = unpack(x[half]);
            z[half] = pack(2.f * m.x, -2.f * m.y);
        }
    }
}

// ----------------------------------------------------------------------------
// Real-input transforms of N = 8192*W points (W = 2, 4, 8) on k_big_nat: the
// same even/odd split as k_r2c / k_c2r (see there) over L = 4096*W complex
// points, one workgroup of W waves per signal.  Wave w handles the pairs
// k = 64*(j*W + w) + lane, j < 32; W_N^k = wn[64w + lane] * wn[64W]^j (wn =
// W_N^0..64W).
//
// k_r2c_big<W>: the k_big_nat pass with the final results kept in registers
// (big_phase2_regs), staged in natural order over the whole LDS (4096*W
// words, lane-contiguous 16-byte writes), then the untangle.
template <int W>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_r2c_big(
    const u32* __restrict__ in, u32* __restrict__ out, const float2* __restrict__ w4096,
    const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo, u32 wn_lo_bits,
    const float2* __restrict__ wn, StartSched sched) {
    extern __shared__ u32 lds[];
    constexpr u32 L = 4096u * W;
    start_delay(sched);
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 w = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 sig = blockIdx.x;
    BigTw bt;
    big_tw_load<W, false>(bt, lane, (int)w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, EpiParams{});
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
    {
        u32 v[16][4];
        const u32* src = in + (size_t)sig * L + 4096u * w + 4u * lane;
#pragma unroll
        for (int t = 0; t < 16; ++t) load4_nt(v[t], src + 256u * t);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 16; ++t) store4(lds + nat_big_word<W>(4096u * w + 256u * t + 4u * lane), v[t]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            d[t][j] = lds[nat_big_word<W>(w + W * ((u32)t + 16u * c16 + 256u * (4u * g + j)))];
    __syncthreads();
    const Dft16Op op = make_dft16(lane, -1);
    u32 e[16][4];
    big_phase1<-1>(d, e, bt, op);
    u32 o[16][4];
    big_phase2_regs<W, BigCfg<W>::ROUNDS, -1>(e, (int)w, lds, o);
    const float2 wb = wn[64u * w + (u32)lane], ws = wn[64u * W];
    __syncthreads();  // every wave is done with the exchange
    constexpr int QN = BigCfg<W>::COLS / (256 * W);
#pragma unroll
    for (int r = 0; r < BigCfg<W>::ROUNDS; ++r)
#pragma unroll
        for (int q = 0; q < QN; ++q)
#pragma unroll
            for (int s = 0; s < W; ++s)
                store4(lds + 2048u * r + 4u * lane + 256u * (q + QN * w) + 4096u * s, o[(r * QN + q) * W + s]);
    __syncthreads();
    u32* const dst = out + (size_t)sig * (L + 1u);
    const cf step = {ws.x, ws.y};
    cf tw = {0.5f * wb.x, 0.5f * wb.y};  // W_N^k / 2
#pragma unroll 8
    for (u32 j = 0; j < 32u; ++j) {
        const u32 k = 64u * (j * W + w) + (u32)lane;  // < L/2
        const cf a = unpack(lds[k]);
        const cf b = unpack(lds[(L - k) & (L - 1u)]);
        const cf hs = {0.5f * (a.x + b.x), 0.5f * (a.y - b.y)};
        const cf dif = {a.x - b.x, a.y + b.y};
        const cf q = cmul(tw, dif);
        dst[k] = pack(hs.x + q.y, hs.y - q.x);
        dst[L - k] = pack(hs.x - q.y, -(hs.y + q.x));
        tw = cmul(tw, step);
    }
    if (w == 0 && lane == 0) {
        const cf a = unpack(lds[L / 2u]);
        dst[L / 2u] = pack(a.x, -a.y);
    }
}

// k_c2r_big<W>: Z' formed pairwise from the bins into the LDS staging of
// k_big_nat (its swizzle), then the k_big_nat inverse pass to the real output.
template <int W>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_c2r_big(
    const u32* __restrict__ in, u32* __restrict__ out, const float2* __restrict__ w4096,
    const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo, u32 wn_lo_bits,
    const float2* __restrict__ wn, StartSched sched) {
    extern __shared__ u32 lds[];
    constexpr u32 L = 4096u * W;
    start_delay(sched);
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    const u32 w = (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 sig = blockIdx.x;
    BigTw bt;
    big_tw_load<W, false>(bt, lane, (int)w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, EpiParams{});
    const float2 wb = wn[64u * w + (u32)lane], ws = wn[64u * W];
    __builtin_amdgcn_sched_barrier(0);
    {
        const u32* const src = in + (size_t)sig * (L + 1u);
        u32 va[32], vb[32];
#pragma unroll
        for (u32 j = 0; j < 32u; ++j) {
            const u32 k = 64u * (j * W + w) + (u32)lane;
            va[j] = __builtin_nontemporal_load(src + k);
            vb[j] = __builtin_nontemporal_load(src + (L - k));
        }
        const u32 vm = (w == 0 && lane == 0) ? src[L / 2u] : 0u;
        const cf step = {ws.x, -ws.y};
        cf tw = {wb.x, -wb.y};  // W_N^-k
#pragma unroll
        for (u32 j = 0; j < 32u; ++j) {
            const u32 k = 64u * (j * W + w) + (u32)lane;
            cf a = unpack(va[j]), b = unpack(vb[j]);
            if (k == 0) a.y = b.y = 0.f;  // bins 0 and L: real parts only
            const cf sum = {a.x + b.x, a.y - b.y};
            const cf dif = {a.x - b.x, a.y + b.y};
            const cf q = cmul(tw, dif);
            lds[nat_big_word<W>(k)] = pack(sum.x - q.y, sum.y + q.x);
            if (k != 0) lds[nat_big_word<W>(L - k)] = pack(sum.x + q.y, q.x - sum.y);
            tw = cmul(tw, step);
        }
        if (w == 0 && lane == 0) {
            const cf a = unpack(vm);
            lds[nat_big_word<W>(L / 2u)] = pack(2.f * a.x, -2.f * a.y);
        }
    }
    __syncthreads();
    u32 d[16][4];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            d[t][j] = lds[nat_big_word<W>(w + W * ((u32)t + 16u * c16 + 256u * (4u * g + j)))];
    __syncthreads();
    const Dft16Op op = make_dft16(lane, +1);
    u32 e[16][4];
    big_phase1<+1>(d, e, bt, op);
    big_phase2<W, BigCfg<W>::ROUNDS, false, +1>(e, (int)w, lds, out + (size_t)sig * L, bt, op, true);
}

// ----------------------------------------------------------------------------
// Host-side launchers (called by the pass driver).
// One-time set-up of a kernel, cached by the kernel's address (every
// template instance of a kernel has the same pointer type, so a static per
// type would be shared by all of them): raise its dynamic-LDS limit to
// `lds` bytes and count the workgroups of `threads` threads resident at once.
struct KernelInfo {
    hipError_t lds_status;
    u32 resident;  // workgroups of this shape resident at once
    u32 cus;       // compute units of the device
};

// Occupancy of `kernel` launched with `threads` threads and `lds` bytes of
// dynamic LDS, cached per launch shape (a kernel may run at several
// workgroup sizes).  The first query of a shape also raises the kernel's
// dynamic LDS limit to `lds` if it is below (never lowers it: another shape
// of the same kernel may need more).
template <class K>
static KernelInfo kernel_info(K kernel, int threads, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, KernelInfo> cache;
    static std::map<const void*, size_t> lds_limit;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple((const void*)kernel, threads, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    KernelInfo info{hipSuccess, 0, 0};
    size_t& limit = lds_limit[(const void*)kernel];
    if (lds > limit) {
        info.lds_status = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (info.lds_status == hipSuccess) limit = lds;
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) {
        info.cus = (u32)cus;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kernel, threads, lds) == hipSuccess)
            info.resident = (u32)(cus * per_cu);
    }
    cache.emplace(key, info);
    return info;
}

// Waves per workgroup of the kernels whose waves each own a 16 KiB LDS slice
// (k_rows_nat, k_r2c, k_c2r): four, or one when the grid has fewer than four
// waves per CU, where four-wave workgroups would leave most CUs idle
// (1024 x 1024 natural order: 64 workgroups on 256 CUs).
template <class K>
static u32 slice_wg(K kernel, u32 waves) {
    const KernelInfo k4 = kernel_info(kernel, 256, 4 * 4096 * sizeof(u32));
    return waves >= 4u * k4.cus ? 4u : 1u;
}

template <int T, int SIGN, bool CACHED, int NT = 16>
static hipError_t launch_rows_t(u32* buf, u32 n_tiles, const float2* w4096, const EpiParams* epi, StartSched sched,
                                u32 wg, hipStream_t st) {
    const u32 waves = (n_tiles + NT - 1) / NT;
    const u32 blocks = (waves + wg - 1) / wg;
    if (sched.groups == 0) sched.delay = 0;
    if (epi) {
        if constexpr (NT != 16) {
            return hipErrorInvalidValue;  // the epilogue pass runs 16 tiles per wave only
        } else {
            auto k = k_rows<T, true, SIGN, CACHED>;
            if (sched.delay) sched.resident = kernel_info(k, (int)(64 * wg), 0).resident;
            hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * wg), 0, st, buf, n_tiles, w4096, *epi, sched);
        }
    } else {
        auto k = k_rows<T, false, SIGN, CACHED, NT>;
        if (sched.delay) sched.resident = kernel_info(k, (int)(64 * wg), 0).resident;
        hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * wg), 0, st, buf, n_tiles, w4096, EpiParams{}, sched);
    }
    return hipGetLastError();
}

// k_rows_half<WPL> over n_tiles / 16 lines: for WPL = 2, wg = 4 puts two
// lines in a workgroup, 1 or 2 one line; for WPL = 4 one line.
template <int SIGN, bool CACHED, int WPL>
static hipError_t launch_rows_half_t(u32* buf, u32 n_tiles, const float2* w4096, StartSched sched, u32 wg,
                                     hipStream_t st) {
    const u32 lines = n_tiles / 16u, lpw = WPL == 2 && wg == 4u ? 2u : 1u;
    const size_t lds = (size_t)lpw * 4096u * sizeof(u32);
    auto k = k_rows_half<SIGN, CACHED, WPL>;
    const KernelInfo ki = kernel_info(k, (int)(64 * WPL * lpw), lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3((lines + lpw - 1u) / lpw), dim3(64 * WPL * lpw), lds, st, buf, lines, w4096, sched);
    return hipGetLastError();
}

template <int WPL>
static hipError_t launch_rows_half_w(u32* buf, u32 n_tiles, int sign, const float2* w4096, StartSched sched, u32 wg,
                                     hipStream_t st) {
    if (sign < 0)
        return sched.cached_out ? launch_rows_half_t<-1, true, WPL>(buf, n_tiles, w4096, sched, wg, st)
                                : launch_rows_half_t<-1, false, WPL>(buf, n_tiles, w4096, sched, wg, st);
    return sched.cached_out ? launch_rows_half_t<+1, true, WPL>(buf, n_tiles, w4096, sched, wg, st)
                            : launch_rows_half_t<+1, false, WPL>(buf, n_tiles, w4096, sched, wg, st);
}

// Only the 4096-point lines (T = 16) have non-temporal output stores to
// switch; shorter lines always store through the caches.  Tiles per wave:
// 16, or 8 / 4 (at least T) for lines of up to 2048 points without an
// epilogue, which puts more waves on a small grid; for 4096-point lines 8
// or 4 tiles per wave is k_rows_half, two or four waves per line.
template <int T>
static hipError_t launch_rows_t(u32* buf, u32 n_tiles, int sign, const float2* w4096, const EpiParams* epi,
                                StartSched sched, u32 wg, u32 tiles, hipStream_t st) {
    if constexpr (T == 16) {
        if (tiles == 8 || tiles == 4) {
            if (epi) return hipErrorInvalidValue;  // the epilogue pass runs 16 tiles per wave only
            return tiles == 8 ? launch_rows_half_w<2>(buf, n_tiles, sign, w4096, sched, wg, st)
                              : launch_rows_half_w<4>(buf, n_tiles, sign, w4096, sched, wg, st);
        }
        if (tiles != 16) return hipErrorInvalidValue;
        if (sched.cached_out)
            return sign < 0 ? launch_rows_t<T, -1, true>(buf, n_tiles, w4096, epi, sched, wg, st)
                            : launch_rows_t<T, +1, true>(buf, n_tiles, w4096, epi, sched, wg, st);
    } else {
        if (tiles == 8)
            return sign < 0 ? launch_rows_t<T, -1, false, 8>(buf, n_tiles, w4096, epi, sched, wg, st)
                            : launch_rows_t<T, +1, false, 8>(buf, n_tiles, w4096, epi, sched, wg, st);
        if constexpr (T <= 4) {
            if (tiles == 4)
                return sign < 0 ? launch_rows_t<T, -1, false, 4>(buf, n_tiles, w4096, epi, sched, wg, st)
                                : launch_rows_t<T, +1, false, 4>(buf, n_tiles, w4096, epi, sched, wg, st);
        }
        if (tiles != 16) return hipErrorInvalidValue;
    }
    return sign < 0 ? launch_rows_t<T, -1, false>(buf, n_tiles, w4096, epi, sched, wg, st)
                    : launch_rows_t<T, +1, false>(buf, n_tiles, w4096, epi, sched, wg, st);
}

hipError_t launch_rows(int T, u32* buf, u32 n_tiles, int sign, const float2* w4096, const EpiParams* epi,
                       StartSched sched, u32 wg_waves, u32 tiles_per_wave, hipStream_t st) {
    if (wg_waves != 1 && wg_waves != 2 && wg_waves != 4) return hipErrorInvalidValue;
    switch (T) {
    case 1: return launch_rows_t<1>(buf, n_tiles, sign, w4096, epi, sched, wg_waves, tiles_per_wave, st);
    case 2: return launch_rows_t<2>(buf, n_tiles, sign, w4096, epi, sched, wg_waves, tiles_per_wave, st);
    case 4: return launch_rows_t<4>(buf, n_tiles, sign, w4096, epi, sched, wg_waves, tiles_per_wave, st);
    case 8: return launch_rows_t<8>(buf, n_tiles, sign, w4096, epi, sched, wg_waves, tiles_per_wave, st);
    case 16: return launch_rows_t<16>(buf, n_tiles, sign, w4096, epi, sched, wg_waves, tiles_per_wave, st);
    default: return hipErrorInvalidValue;
    }
}

template <int T>
static hipError_t launch_cols_t(u32* buf, u32 s_log2, u32 outer, int sign, const float2* w4096,
                                const EpiParams* epi, hipStream_t st) {
    if constexpr (T <= 4) {
        if (!epi && s_log2 >= 6) {
            const u32 waves = outer * ((1u << s_log2) >> 6);
            hipLaunchKernelGGL((k_cols_q<T>), dim3((waves + 3) / 4), dim3(256), 0, st, buf, s_log2, waves, sign, w4096);
            return hipGetLastError();
        }
    }
    if constexpr (T == 8 || T == 16) {
        if (!epi && s_log2 >= 6) {
            constexpr size_t lds = 16u * T * 64u * sizeof(u32);  // 16T rows x 64 columns
            const u32 groups = outer << (s_log2 - 6u);
            auto k = sign < 0 ? k_cols_w<T, -1> : k_cols_w<T, +1>;
            const KernelInfo ki = kernel_info(k, 256, lds);
            if (ki.lds_status != hipSuccess) return ki.lds_status;
            hipLaunchKernelGGL(k, dim3(groups), dim3(256), lds, st, buf, s_log2, groups, w4096);
            return hipGetLastError();
        }
    }
    const u32 waves = outer * ((1u << s_log2) >> 4);
    const u32 blocks = (waves + 3) / 4;
    if (epi) hipLaunchKernelGGL((k_cols<T, true>), dim3(blocks), dim3(256), 0, st, buf, s_log2, waves, sign, w4096, *epi);
    else hipLaunchKernelGGL((k_cols<T, false>), dim3(blocks), dim3(256), 0, st, buf, s_log2, waves, sign, w4096, EpiParams{});
    return hipGetLastError();
}

template <int T, int SIGN>
static hipError_t launch_rows_nat_t(const u32* in, u32* out, u32 n_tiles, const float2* w4096, u32 nx_log2,
                                    const u32* xpos, StartSched sched, hipStream_t st) {
    // cached output stores switch only for 4096-point lines (as launch_rows_t)
    auto k = k_rows_nat<T, SIGN, false>;
    if constexpr (T == 16)
        if (sched.cached_out) k = k_rows_nat<T, SIGN, true>;
    const u32 waves = (n_tiles + 15) / 16;
    const u32 wg = slice_wg(k, waves);
    const size_t lds = wg * 4096 * sizeof(u32);  // 16 KiB per wave
    const KernelInfo ki = kernel_info(k, (int)(64 * wg), lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3((waves + wg - 1) / wg), dim3(64 * wg), lds, st, in, out, n_tiles, w4096, nx_log2,
                       xpos, sched);
    return hipGetLastError();
}

// k_rows_half_nat<WPL> over n_tiles / 16 lines, one line (WPL waves,
// 32 KiB of LDS) per workgroup.
template <int SIGN, bool CACHED, int WPL>
static hipError_t launch_rows_half_nat_t(const u32* in, u32* out, u32 n_tiles, const float2* w4096, u32 nx_log2,
                                         const u32* xpos, StartSched sched, hipStream_t st) {
    const u32 lines = n_tiles / 16u;
    const size_t lds = 8192u * sizeof(u32);
    auto k = k_rows_half_nat<SIGN, CACHED, WPL>;
    const KernelInfo ki = kernel_info(k, 64 * WPL, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3(lines), dim3(64 * WPL), lds, st, in, out, lines, w4096, nx_log2, xpos, sched);
    return hipGetLastError();
}

template <int WPL>
static hipError_t launch_rows_half_nat_w(const u32* in, u32* out, u32 n_tiles, int sign, const float2* w4096,
                                         u32 nx_log2, const u32* xpos, StartSched sched, hipStream_t st) {
    if (sign < 0)
        return sched.cached_out ? launch_rows_half_nat_t<-1, true, WPL>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st)
                                : launch_rows_half_nat_t<-1, false, WPL>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    return sched.cached_out ? launch_rows_half_nat_t<+1, true, WPL>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st)
                            : launch_rows_half_nat_t<+1, false, WPL>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
}

hipError_t launch_rows_nat(int T, const u32* in, u32* out, u32 n_tiles, int sign, const float2* w4096, u32 nx_log2,
                           const u32* xpos, StartSched sched, u32 tiles_per_wave, hipStream_t st) {
    if (nx_log2 && !xpos) return hipErrorInvalidValue;
    if ((tiles_per_wave == 8 || tiles_per_wave == 4) && T == 16)
        return tiles_per_wave == 8
                   ? launch_rows_half_nat_w<2>(in, out, n_tiles, sign, w4096, nx_log2, xpos, sched, st)
                   : launch_rows_half_nat_w<4>(in, out, n_tiles, sign, w4096, nx_log2, xpos, sched, st);
    if (tiles_per_wave != 16) return hipErrorInvalidValue;
    switch (T * 2 + (sign < 0 ? 0 : 1)) {
    case 2: return launch_rows_nat_t<1, -1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 3: return launch_rows_nat_t<1, +1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 4: return launch_rows_nat_t<2, -1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 5: return launch_rows_nat_t<2, +1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 8: return launch_rows_nat_t<4, -1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 9: return launch_rows_nat_t<4, +1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 16: return launch_rows_nat_t<8, -1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 17: return launch_rows_nat_t<8, +1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 32: return launch_rows_nat_t<16, -1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    case 33: return launch_rows_nat_t<16, +1>(in, out, n_tiles, w4096, nx_log2, xpos, sched, st);
    default: return hipErrorInvalidValue;
    }
}

template <int T>
static hipError_t launch_r2c_t(int sign, const u32* in, u32* out, u32 n_tiles, const float2* w4096,
                               const float2* wn, hipStream_t st) {
    auto k = sign < 0 ? k_r2c<T> : k_c2r<T>;
    const u32 waves = (n_tiles + 15) / 16;
    const u32 wg = slice_wg(k, waves);
    const size_t lds = wg * 4096 * sizeof(u32);  // 16 KiB per wave
    const KernelInfo ki = kernel_info(k, (int)(64 * wg), lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    hipLaunchKernelGGL(k, dim3((waves + wg - 1) / wg), dim3(64 * wg), lds, st, in, out, n_tiles, w4096, wn);
    return hipGetLastError();
}

template <int W>
static hipError_t launch_r2c_big_t(int sign, const u32* in, u32* out, u32 n_signals, const float2* w4096,
                                   const float2* wn_hi, const float2* wn_lo, u32 lo_bits, const float2* wn,
                                   StartSched sched, hipStream_t st) {
    const size_t lds = 4096u * W * sizeof(u32);  // the whole half-size signal
    static_assert(4096u * W >= (size_t)BigCfg<W>::LDS_WORDS, "LDS covers the exchange");
    auto k = sign < 0 ? k_r2c_big<W> : k_c2r_big<W>;
    const KernelInfo ki = kernel_info(k, 64 * W, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, in, out, w4096, wn_hi, wn_lo, lo_bits, wn, sched);
    return hipGetLastError();
}

hipError_t launch_r2c_big(int W, int sign, const u32* in, u32* out, u32 n_signals, const float2* w4096,
                          const float2* wn_hi, const float2* wn_lo, u32 lo_bits, const float2* wn, StartSched sched,
                          hipStream_t st) {
    switch (W) {
    case 2: return launch_r2c_big_t<2>(sign, in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, wn, sched, st);
    case 4: return launch_r2c_big_t<4>(sign, in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, wn, sched, st);
    case 8: return launch_r2c_big_t<8>(sign, in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, wn, sched, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_real_pairs(int sign, const u32* in, u32* out, u32 l_log2, u32 n_signals, const float2* hi,
                             const float2* lo, u32 lo_bits, hipStream_t st) {
    if (l_log2 < 2 || l_log2 > 26) return hipErrorInvalidValue;
    const u32 pairs = n_signals << (l_log2 - 1u);
    const u32 blocks = (pairs + 255u) / 256u;
    if (sign < 0) hipLaunchKernelGGL(k_pairs<-1>, dim3(blocks), dim3(256), 0, st, in, out, l_log2, pairs, hi, lo, lo_bits);
    else hipLaunchKernelGGL(k_pairs<+1>, dim3(blocks), dim3(256), 0, st, in, out, l_log2, pairs, hi, lo, lo_bits);
    return hipGetLastError();
}

hipError_t launch_r2c(int T, int sign, const u32* in, u32* out, u32 n_tiles, const float2* w4096, const float2* wn,
                      hipStream_t st) {
    switch (T) {
    case 1: return launch_r2c_t<1>(sign, in, out, n_tiles, w4096, wn, st);
    case 2: return launch_r2c_t<2>(sign, in, out, n_tiles, w4096, wn, st);
    case 4: return launch_r2c_t<4>(sign, in, out, n_tiles, w4096, wn, st);
    case 8: return launch_r2c_t<8>(sign, in, out, n_tiles, w4096, wn, st);
    case 16: return launch_r2c_t<16>(sign, in, out, n_tiles, w4096, wn, st);
    default: return hipErrorInvalidValue;
    }
}

template <int W, int SIGN>
static hipError_t launch_big_nat_t(const u32* in, u32* out, u32 n_signals, const float2* w4096, const float2* wn_hi,
                                   const float2* wn_lo, u32 lo_bits, u32 nx_log2, const RevSpec& x_pos,
                                   StartSched sched, hipStream_t st) {
    // the whole signal (W < 16) or half of it (W = 16), >= the exchange's LDS
    constexpr size_t words = W < 16 ? 4096u * W : 32768u;
    static_assert(words >= (size_t)BigCfg<W>::LDS_WORDS, "LDS covers the exchange");
    const size_t lds = words * sizeof(u32);
    auto k = sched.cached_out ? k_big_nat<W, SIGN, true> : k_big_nat<W, SIGN, false>;
    const KernelInfo ki = kernel_info(k, 64 * W, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, in, out, w4096, wn_hi, wn_lo, lo_bits, nx_log2,
                       x_pos, sched);
    return hipGetLastError();
}

hipError_t launch_big_nat(int W, const u32* in, u32* out, u32 n_signals, int sign, const float2* w4096,
                          const float2* wn_hi, const float2* wn_lo, u32 lo_bits, u32 nx_log2, const RevSpec& x_pos,
                          StartSched sched, hipStream_t st) {
    switch (W * 2 + (sign < 0 ? 0 : 1)) {
    case 4: return launch_big_nat_t<2, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 5: return launch_big_nat_t<2, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 8: return launch_big_nat_t<4, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 9: return launch_big_nat_t<4, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 16: return launch_big_nat_t<8, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 17: return launch_big_nat_t<8, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 32: return launch_big_nat_t<16, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    case 33: return launch_big_nat_t<16, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, nx_log2, x_pos, sched, st);
    default: return hipErrorInvalidValue;
    }
}

template <int W, int SIGN>
static hipError_t launch_cols_big_t(u32* buf, u32 s_log2, u32 outer, const float2* w4096, StartSched sched,
                                    hipStream_t st) {
    if constexpr (W == 2) {
        if (s_log2 >= 5u) {  // 32 columns per workgroup, 16-byte accesses
            const size_t lds = ColsBigWCfg<W>::LDS_WORDS * sizeof(u32);
            auto k = k_cols_bigw<W, SIGN>;
            const KernelInfo ki = kernel_info(k, 128 * W, lds);
            if (ki.lds_status != hipSuccess) return ki.lds_status;
            if (sched.groups == 0) sched.delay = 0;
            if (sched.delay) sched.resident = ki.resident;
            const u32 n_groups = outer << (s_log2 - 5u);
            hipLaunchKernelGGL(k, dim3(n_groups), dim3(128 * W), lds, st, buf, s_log2, n_groups, w4096, sched);
            return hipGetLastError();
        }
    }
    const size_t lds = ColsBigCfg<W>::LDS_WORDS * sizeof(u32);
    auto k = k_cols_big<W, SIGN>;
    const KernelInfo ki = kernel_info(k, 64 * W, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    const u32 n_groups = outer << (s_log2 - 4u);
    hipLaunchKernelGGL(k, dim3(n_groups), dim3(64 * W), lds, st, buf, s_log2, n_groups, w4096, sched);
    return hipGetLastError();
}

hipError_t launch_cols_big(int W, u32* buf, u32 s_log2, u32 outer, int sign, const float2* w4096, StartSched sched,
                           hipStream_t st) {
    if (s_log2 < 4u || ((size_t)outer << (s_log2 - 4u)) > 0x7fffffffu) return hipErrorInvalidValue;
    switch (W * 2 + (sign < 0 ? 0 : 1)) {
    case 4: return launch_cols_big_t<2, -1>(buf, s_log2, outer, w4096, sched, st);
    case 5: return launch_cols_big_t<2, +1>(buf, s_log2, outer, w4096, sched, st);
    case 8: return launch_cols_big_t<4, -1>(buf, s_log2, outer, w4096, sched, st);
    case 9: return launch_cols_big_t<4, +1>(buf, s_log2, outer, w4096, sched, st);
    case 16: return launch_cols_big_t<8, -1>(buf, s_log2, outer, w4096, sched, st);
    case 17: return launch_cols_big_t<8, +1>(buf, s_log2, outer, w4096, sched, st);
    default: return hipErrorInvalidValue;
    }
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

template <int W, int SIGN>
static hipError_t launch_big_t(u32* buf, u32 n_signals, const float2* w4096, const float2* wn_hi,
                               const float2* wn_lo, u32 lo_bits, const EpiParams* epi, StartSched sched,
                               hipStream_t st) {
    const size_t lds = BigCfg<W>::LDS_WORDS * sizeof(u32);
    if (sched.groups == 0) sched.delay = 0;
    if (epi) {
        // the folded epilogue covers a first pass of 65536 contiguous points
        // followed by a pass at step 65536 along the same axis
        if constexpr (W == 16) {
            if (epi->shift != 0 || epi->s_log2 != 16) return hipErrorInvalidValue;
            auto k = sched.cached_out ? k_big<W, true, SIGN, true> : k_big<W, true, SIGN, false>;
            const KernelInfo ki = kernel_info(k, 64 * W, lds);
            if (ki.lds_status != hipSuccess) return ki.lds_status;
            if (sched.delay) sched.resident = ki.resident;
            hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, buf, w4096, wn_hi, wn_lo, lo_bits, *epi,
                               sched);
        } else {
            return hipErrorInvalidValue;
        }
    } else {
        auto k = sched.cached_out ? k_big<W, false, SIGN, true> : k_big<W, false, SIGN, false>;
        const KernelInfo ki = kernel_info(k, 64 * W, lds);
        if (ki.lds_status != hipSuccess) return ki.lds_status;
        if (sched.delay) sched.resident = ki.resident;
        hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, buf, w4096, wn_hi, wn_lo, lo_bits,
                           EpiParams{}, sched);
    }
    return hipGetLastError();
}

template <int W>
static hipError_t launch_big_t(u32* buf, u32 n_signals, int sign, const float2* w4096, const float2* wn_hi,
                               const float2* wn_lo, u32 lo_bits, const EpiParams* epi, StartSched sched,
                               hipStream_t st) {
    return sign < 0 ? launch_big_t<W, -1>(buf, n_signals, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st)
                    : launch_big_t<W, +1>(buf, n_signals, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st);
}

hipError_t launch_big(int W, u32* buf, u32 n_signals, int sign, const float2* w4096, const float2* wn_hi,
                      const float2* wn_lo, u32 lo_bits, const EpiParams* epi, StartSched sched, hipStream_t st) {
    switch (W) {
    case 2: return launch_big_t<2>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st);
    case 4: return launch_big_t<4>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st);
    case 8: return launch_big_t<8>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st);
    case 16: return launch_big_t<16>(buf, n_signals, sign, w4096, wn_hi, wn_lo, lo_bits, epi, sched, st);
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
    // wave-uniform (scalar) so that the per-line offsets are scalar loads;
    // lane_id() keeps the lane offset inside the task loop (hoisted out of it,
    // it stays live across every size class and the kernel spills)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id();
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
                                                const float2* __restrict__ w4096, MultiTables tb,
                                                StartSched sched) {
    extern __shared__ u32 lds[];
    start_delay(sched);
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

// ----------------------------------------------------------------------------
// Natural-order uneven batch (k_multi_nat): the tasks of k_multi on
// natural-order input, in -> out (in place allowed: each signal belongs to
// one wave or group, which reads it completely before its first store).
// Row classes of 256..2048 points pass each wave's lines through the wave's
// own 8 KiB of LDS (words k*L + natural index for the k-th line of a call)
// and read them back in the digit-reversed register layout (nat_rows_read's
// formula); a 4096-point line passes in two rounds split by index bit 8.
// Big classes stage every group's signal in two rounds split by index bit
// 9 or 11 (W * 8 KiB per group per round, 128 KiB in all).

template <int T, int SIGN>
__device__ __forceinline__ void multi_rows_nat(const u32* in, u32* out, const MultiTask& tk,
                                               const u32* __restrict__ offsets, const float2* w4096, u32* lds) {
    constexpr int LPW = 16 / T;
    constexpr u32 L = 256u * T;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = lane_id();
    const int first_line = wave * LPW;
    const int lines = min((int)tk.count - first_line, LPW);
    const int c16 = lane & 15, g = lane >> 4;
    const u32 lane_off = 16u * c16 + 4u * g;
    const u32* offs = offsets + tk.first + first_line;
    if constexpr (T == 16) {
        if (lines <= 0) return;
        // one 4096-point line per wave through the wave's 2048-word slice in
        // two rounds: round r holds the natural indices with bit 8 = r, i.e.
        // the loaded tiles with t & 1 = r and the read words with j & 1 = r
        // (reader index t + 16*c16 + 256*(4g + j)), so that which registers
        // each round writes and reads is fixed at compile time
        u32* const wl = lds + (u32)wave * 2048u;
        auto load = [&](u32 (&d)[16][4]) {
            u32 v[16][4];
            const u32* src = in + offs[0] + 4u * lane;
#pragma unroll
            for (int t = 0; t < 16; ++t) load4_nt(v[t], src + 256u * t);
            // reader word of (t, j): base_j + (t ^ xlo_j) (nat_split_lane)
            u32 base[4], xlo[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) nat_split_lane<8, 4>(16u * c16 + 256u * (4u * g + j), base[j], xlo[j]);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
#pragma unroll
                for (int tt = 0; tt < 8; ++tt)
                    store4(wl + nat_split_word<8>(256u * (2u * tt + r) + 4u * lane), v[2 * tt + r]);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
                for (int t = 0; t < 16; ++t)
#pragma unroll
                    for (int j = r; j < 4; j += 2) d[t][j] = wl[base[j] + ((u32)t ^ xlo[j])];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        };
        rows_body<16, false, SIGN>([&](int t) { return out + offs[0] + 256u * t + lane_off; }, 16u, 0u, w4096,
                                   EpiParams{}, load);
    } else {
        if (lines <= 0) return;
        u32* const wl = lds + (u32)wave * 2048u;  // 8 KiB per wave: NT * 256 words
        constexpr int NT = (T == 1 || T == 8) ? 8 : 4;  // as multi_rows
        for (int t0 = 0; t0 < lines * T; t0 += NT) {
            const u32 ntl = (u32)min(lines * T - t0, NT);
            auto load = [&](u32 (&d)[NT][4]) {
#pragma unroll
                for (int k = 0; k < NT / T; ++k)
                    if ((u32)(k * T) < ntl) {
                        const u32* src = in + offs[t0 / T + k] + 4u * lane;
#pragma unroll
                        for (int tt = 0; tt < T; ++tt) {
                            u32 v[4];
                            load4_nt(v, src + 256u * tt);
                            store4(wl + nat_lds_word((u32)k * L + 256u * tt + 4u * lane), v);
                        }
                    }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                nat_rows_read<T, NT>(d, wl, ntl, lane);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            };
            rows_body<T, false, SIGN, NT>(
                [&](int t) { return out + offs[(t0 + t) / T] + 256u * ((t0 + t) % T) + lane_off; }, ntl, 0u, w4096,
                EpiParams{}, load);
        }
    }
}

template <int W, int SIGN>
__device__ __forceinline__ void multi_big_nat(const u32* in, u32* out, const MultiTask& tk,
                                              const u32* __restrict__ offsets, const float2* w4096,
                                              const MultiTables& tb, u32* lds) {
    constexpr int M = W == 2 ? 13 : W == 4 ? 14 : W == 8 ? 15 : 16;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int grp = wave / W, w = wave % W;
    const int lane = lane_id();
    const int c16 = lane & 15, g = lane >> 4;
    // a group without a signal (last task of a class) reads the task's first
    // signal, so that every load stays in bounds, and stores nothing
    const bool active = grp < (int)tk.count;
    const u32 off = offsets[tk.first + (active ? grp : 0)];
    u32 v[16][4];
    {
        const u32* src = in + off + 4096u * (u32)w + 4u * lane;
#pragma unroll
        for (int t = 0; t < 16; ++t) load4_nt(v[t], src + 256u * t);
    }
    // Two rounds over every group at once: round r stages the natural
    // indices n with bit B of n = r at word nat_split_word<B>(n) of the
    // group's W * 2048 words.  B is chosen so that the writer's half is a
    // set of its loaded tiles (n = 4096w + 256t + ...: tile bit B - 8) and,
    // for W = 2, 4, 8, the reader's half a set of its registers (reader
    // n = w + W*(t + 16*c16 + 256*(4g + j)): bit 9 = j & 1 for W = 2, bit
    // 11 = j >> 1 for W = 4 and j & 1 for W = 8); for W = 16 bit 11 is
    // c16 >> 3, a lane predicate.
    constexpr int B = W == 2 ? 9 : 11;
    constexpr int LW = W == 2 ? 5 : W == 4 ? 6 : W == 8 ? 7 : 8;  // log2(16W)
    // reader word of (t, j): base_j + ((w + W t) ^ xlo_j) (nat_split_lane)
    u32 base[4], xlo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        nat_split_lane<B, LW>(16u * W * (u32)c16 + 256u * W * (4u * g + (u32)j), base[j], xlo[j]);
    u32 d[16][4];
    u32* const region = lds + (u32)grp * (2048u * W);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (r) __syncthreads();  // round 0's reads are done
#pragma unroll
        for (int t = 0; t < 16; ++t)
            if (((t >> (B - 8)) & 1) == r)
                store4(region + nat_split_word<B>(4096u * (u32)w + 256u * t + 4u * lane), v[t]);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const u32 word = base[j] + (((u32)w + W * (u32)t) ^ xlo[j]);
                if constexpr (W == 16) {
                    if ((c16 >> 3) == r) d[t][j] = region[word];
                } else {
                    if ((W == 4 ? (j >> 1) : (j & 1)) == r) d[t][j] = region[word];
                }
            }
    }
    __syncthreads();  // every wave has read its data before the exchange reuses the LDS
    BigTw bt;
    big_tw_load<W, false>(bt, lane_id(), w, 0u, w4096, tb.hi[M - 13], tb.lo[M - 13], (u32)(M / 2), EpiParams{});
    const Dft16Op op = make_dft16(lane_id(), SIGN);
    u32 e[16][4];
    big_phase1<SIGN>(d, e, bt, op);
    big_phase2<W, 2, false, SIGN>(e, w, lds + grp * (W * 2048), out + off, bt, op, active);
}

template <int SIGN>
__global__ __launch_bounds__(1024) void k_multi_nat(const u32* in, u32* out, const MultiTask* __restrict__ tasks,
                                                    u32 n_tasks, const u32* __restrict__ offsets,
                                                    const float2* __restrict__ w4096, MultiTables tb,
                                                    StartSched sched) {
    extern __shared__ u32 lds[];
    start_delay(sched);
    for (u32 ti = blockIdx.x; ti < n_tasks; ti += gridDim.x) {
        const MultiTask tk = tasks[ti];
        switch (tk.size_log2) {
        case 8: multi_rows_nat<1, SIGN>(in, out, tk, offsets, w4096, lds); break;
        case 9: multi_rows_nat<2, SIGN>(in, out, tk, offsets, w4096, lds); break;
        case 10: multi_rows_nat<4, SIGN>(in, out, tk, offsets, w4096, lds); break;
        case 11: multi_rows_nat<8, SIGN>(in, out, tk, offsets, w4096, lds); break;
        case 12: multi_rows_nat<16, SIGN>(in, out, tk, offsets, w4096, lds); break;
        case 13: multi_big_nat<2, SIGN>(in, out, tk, offsets, w4096, tb, lds); break;
        case 14: multi_big_nat<4, SIGN>(in, out, tk, offsets, w4096, tb, lds); break;
        case 15: multi_big_nat<8, SIGN>(in, out, tk, offsets, w4096, tb, lds); break;
        case 16: multi_big_nat<16, SIGN>(in, out, tk, offsets, w4096, tb, lds); break;
        default: break;
        }
        __syncthreads();  // the next task reuses the LDS
    }
}

template <int SIGN>
static hipError_t launch_multi_nat_t(const u32* in, u32* out, const MultiTask* tasks, u32 n_tasks,
                                     const u32* offsets, u32 grid, const float2* w4096, const MultiTables& tb,
                                     StartSched sched, hipStream_t st) {
    constexpr size_t lds = 16 * 2048 * sizeof(u32);  // 128 KiB
    auto k = k_multi_nat<SIGN>;
    const KernelInfo ki = kernel_info(k, 1024, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1024), lds, st, in, out, tasks, n_tasks, offsets, w4096, tb, sched);
    return hipGetLastError();
}

hipError_t launch_multi_nat(const u32* in, u32* out, const MultiTask* tasks, u32 n_tasks, const u32* offsets,
                            u32 grid, int sign, const float2* w4096, const MultiTables& tb, StartSched sched,
                            hipStream_t st) {
    if (n_tasks == 0) return hipSuccess;
    if (grid == 0 || grid > n_tasks) grid = n_tasks;
    return sign < 0 ? launch_multi_nat_t<-1>(in, out, tasks, n_tasks, offsets, grid, w4096, tb, sched, st)
                    : launch_multi_nat_t<+1>(in, out, tasks, n_tasks, offsets, grid, w4096, tb, sched, st);
}

template <int SIGN>
static hipError_t launch_multi_t(u32* buf, const MultiTask* tasks, u32 n_tasks, const u32* offsets, u32 grid,
                                 const float2* w4096, const MultiTables& tb, StartSched sched, hipStream_t st) {
    constexpr size_t lds = 16 * 2048 * sizeof(u32);  // 128 KiB
    auto k = k_multi<SIGN>;
    const KernelInfo ki = kernel_info(k, 1024, lds);
    if (ki.lds_status != hipSuccess) return ki.lds_status;
    if (sched.groups == 0) sched.delay = 0;
    if (sched.delay) sched.resident = ki.resident;
    hipLaunchKernelGGL(k, dim3(grid), dim3(1024), lds, st, buf, tasks, n_tasks, offsets, w4096, tb, sched);
    return hipGetLastError();
}

hipError_t launch_multi(u32* buf, const MultiTask* tasks, u32 n_tasks, const u32* offsets, u32 grid, int sign,
                        const float2* w4096, const MultiTables& tb, StartSched sched, hipStream_t st) {
    if (n_tasks == 0) return hipSuccess;
    if (grid == 0 || grid > n_tasks) grid = n_tasks;
    return sign < 0 ? launch_multi_t<-1>(buf, tasks, n_tasks, offsets, grid, w4096, tb, sched, st)
                    : launch_multi_t<+1>(buf, tasks, n_tasks, offsets, grid, w4096, tb, sched, st);
}

}  // namespace tcfft
