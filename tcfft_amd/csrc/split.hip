// split.hip -- split-precision FFT: FP32 complex in/out, FP16 hi+lo on MFMA.
//
// Every radix-16 stage computes D = Fh*Xh + Fh*Xl + Fl*Xh with
// v_mfma_f32_16x16x32_f16 and FP32 accumulation, where X = Xh + Xl is the
// FP16 hi/lo split of the scaled FP32 data and F = Fh + Fl that of the
// DFT-16 matrix.  The dropped Fl*Xl term and the two FP16 roundings leave
// about 2^-22 relative error per stage.  Each wave scales its data by one
// power of two so that the largest component is below 2^13 (no overflow in
// FP16; values far below the largest lose only bits that are negligible next
// to it in the L2 norm); the bound carries to the second stage (a DFT-16
// grows magnitudes at most 16-fold, so that stage uses scale/16) and the
// inverse scale is applied to the FP32 results.  Twiddles come from
// FP64-generated tables.
//
// N = 65536 = 256 x 256 as a four-step in two passes over FP32 data, both
// the same kernel: each wave computes the 256-point DFTs down 16 adjacent
// columns of a 256 x 256 matrix, entirely in registers (no LDS):
//   pass 1: x[c + 256 r] -> Y_c[k] = W_N^{c k} sum_r x W_256^{r k},
//           stored transposed as T[k + 256 c] (4 consecutive k per lane,
//           128-byte runs per row);
//   pass 2: T[k + 256 c] -> X[k + 256 m] = sum_c T W_256^{c m}.
// Column DFT-256 over rows r = r0 + 16 r1, output k = k0 + 16 k1:
//   stage 1 (tile r0): B-form over r1 -> rows k0; twiddle W_256^{r0 k0};
//   transpose (g,i) <-> tile; stage 2 (tile k0): B-form over r0 -> rows k1.
#include "fft_device.hpp"
#include "fft_kernels.hpp"

namespace tcfft {

struct Dft16Split {
    Dft16Op h, l;
};

__device__ __forceinline__ Dft16Split make_dft16_split(int lane, int sign) {
    const int f = lane & 15, g = lane >> 4;
    Dft16Split s;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = 4 * g + j;
        const int m = (n * f) & 15;
        const float fr = cos16_exact(m);
        const float fi = sign < 0 ? -sin16_exact(m) : sin16_exact(m);
        const _Float16 frh = (_Float16)fr, fih = (_Float16)fi;
        const _Float16 frl = (_Float16)(fr - (float)frh), fil = (_Float16)(fi - (float)fih);
        s.h.re[2 * j] = frh;
        s.h.re[2 * j + 1] = -fih;
        s.h.im[2 * j] = fih;
        s.h.im[2 * j + 1] = frh;
        s.l.re[2 * j] = frl;
        s.l.re[2 * j + 1] = -fil;
        s.l.im[2 * j] = fil;
        s.l.im[2 * j + 1] = frl;
    }
    return s;
}

// B-form DFT-16 of the four (scaled) points (re[j], im[j]) * s of this lane
// (column lane%16, rows 4g+j): FP32 results for rows 4g+i in (ore, oim).
template <bool SCALE>
__device__ __forceinline__ void dft16_split_b(const float* re, const float* im, float s, const Dft16Split& F,
                                              f32x4& ore, f32x4& oim) {
    u32 hi[4], lo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // hi = fp16(x*s); lo = fp16(x*s - hi) (the difference is exact in
        // FP32); without SCALE, s = 1
        const float xr = SCALE ? re[j] * s : re[j], xi = SCALE ? im[j] * s : im[j];
        hi[j] = pack(xr, xi);
        const cf h = unpack(hi[j]);
        lo[j] = pack(xr - h.x, xi - h.y);
    }
    const f16x8 bh = as_operand(hi[0], hi[1], hi[2], hi[3]);
    const f16x8 bl = as_operand(lo[0], lo[1], lo[2], lo[3]);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    ore = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.re, bh, z, 0, 0, 0);
    ore = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.re, bl, ore, 0, 0, 0);
    ore = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.l.re, bh, ore, 0, 0, 0);
    oim = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.im, bh, z, 0, 0, 0);
    oim = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.im, bl, oim, 0, 0, 0);
    oim = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.l.im, bh, oim, 0, 0, 0);
}

// 2^(13 - e) for the wave-wide largest magnitude m = f * 2^e (f in [0.5, 1)),
// and its inverse; 1 for m = 0.  s = 2^(sb - 127) is clamped to at most
// 2^126 (sb <= 253) so that its inverse 2^(127 - sb), biased exponent
// 254 - sb >= 1, is a normal float and s * inv == 1 exactly: a wave whose
// largest component is below 2^-114 is scaled by 2^126 only (its scaled
// values stay below 2^13 all the same).
__device__ __forceinline__ void wave_scale(float m, float& s, float& inv) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
    const int ex = (int)((__float_as_uint(m) >> 23) & 255u);  // m in [2^(ex-127), 2^(ex-126))
    if (ex == 0) {
        s = inv = 1.f;
        return;
    }
    const int sb = min(max(266 - ex, 1), 253);
    s = __uint_as_float((u32)sb << 23);
    inv = __uint_as_float((u32)(254 - sb) << 23);
}

// src and dst alias in pass 2 (in place on the intermediate), so neither is
// __restrict__: a workgroup reads its 64 columns completely before its first
// store, and workgroups own disjoint columns.
// Start schedule (as the first-pass kernels): the resident workgroups in
// `groups` groups by their index within an XCD, group k sleeping k * delay
// quanta of 1024 cycles, so that loads, compute and stores overlap.
__device__ __forceinline__ void split_start_delay(const StartSched& sched) {
    if (sched.delay && blockIdx.x < sched.resident) {
        const u32 quanta = ((blockIdx.x >> 3) % sched.groups) * sched.delay;
        for (u32 i = 0; i < quanta; ++i) __builtin_amdgcn_s_sleep(16);
    }
}

// One workgroup's share of pass PASS: block blk = 4 * signal + column group
// (64 columns of the signal's 256 x 256 matrix).
template <int PASS, int SIGN>
__device__ __forceinline__ void split_col_body(const float2* src, float2* dst, u32 blk,
                                               const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                               const float2* __restrict__ wn_lo) {
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int c16 = lane & 15, g = lane >> 4;
    const u32 sig = blk >> 2;
    const u32 col = 64u * (blk & 3u) + 16u * w + c16;
    const size_t sbase = (size_t)sig * 65536u;

    // the stage-1 twiddle step W_256^{4g+i} first (global loads complete in
    // issue order, so a table value loaded after the data would wait for it)
    float2 t_step[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t_step[i] = w4096[16 * (4 * g + i)];
    __builtin_amdgcn_sched_barrier(0);
    // rows r = r0 + 16 (4g + j): re/im of tile r0, word j
    float dr[16][4], di[16][4];
    {
        const float2* p = src + sbase + col + 16384u * g;  // row 64g, column col
#pragma unroll
        for (int r0 = 0; r0 < 16; ++r0)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // pass 2 reads the intermediate that pass 1 has just stored
                // through the caches: cached loads; pass 1 streams the input
                const f32x2* a = reinterpret_cast<const f32x2*>(p + 256u * (r0 + 16u * j));
                f32x2 v;
                if constexpr (PASS == 2)
                    v = *a;
                else
                    v = __builtin_nontemporal_load(a);
                dr[r0][j] = v[0];
                di[r0][j] = v[1];
            }
    }
    __builtin_amdgcn_sched_barrier(0);
    float m = 0.f;
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmaxf(m, fmaxf(fabsf(dr[t][j]), fabsf(di[t][j])));
    float s, inv;
    wave_scale(m, s, inv);
    const Dft16Split F = make_dft16_split(lane, SIGN);

    // Stage 1 (tile r0; rows (g,i) = k0) and its epilogue factor
    // W_256^{r0 k0} / 16, the 1/16 keeping the stage-2 inputs within the same
    // bound (a DFT-16 grows magnitudes at most 16-fold).
    cf step[4], rec[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        step[i] = cf{t_step[i].x, sg * t_step[i].y};
        rec[i] = cf{0.0625f, 0.f};
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 ore, oim;
        dft16_split_b<true>(dr[t], di[t], s, F, ore, oim);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const cf v = cmul(cf{ore[i], oim[i]}, rec[i]);
            if (t < 15) rec[i] = cmul(rec[i], step[i]);
            dr[t][i] = v.x;
            di[t][i] = v.y;
        }
    }
    // (g,i) = k0, tile = r0  ->  (g,i) = r0, tile = k0
    {
        float tr[16][4], ti[16][4];
        transpose_gi_tile(dr, tr);
        transpose_gi_tile(di, ti);
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dr[t][i] = tr[t][i];
                di[t][i] = ti[t][i];
            }
    }
    // Stage 2 (tile k0; rows (g,i) = k1), already in scaled units; inv2 undoes
    // the scale s and the 1/16.
    const float inv2 = inv * 16.f;
    if constexpr (PASS == 2) {
        // X[k + 256 m], m = k0 + 16 k1: column col, row m; each tile is stored
        // as soon as it is transformed
        float2* q = dst + sbase + col + 16384u * g;
        // the DFT-16 operand again, from the opaque lane id: kept live from
        // stage 1 instead, it is spilled (12 VGPRs of scratch per lane)
        const Dft16Split F2 = make_dft16_split(lane_id(), SIGN);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 ore, oim;
            dft16_split_b<false>(dr[t], di[t], 1.f, F2, ore, oim);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const f32x2 v = {ore[i] * inv2, oim[i] * inv2};
                __builtin_nontemporal_store(v, reinterpret_cast<f32x2*>(q + 256u * (t + 16u * i)));
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 ore, oim;
            dft16_split_b<false>(dr[t], di[t], 1.f, F, ore, oim);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dr[t][i] = ore[i];
                di[t][i] = oim[i];
            }
        }
        // Y_c[k] W_N^{c k}, k = (4g + i) + 16 t: W_N^{c(4g+i)} * (W_N^{16c})^t
        // (two-level table), with the inverse scale; loaded here, after the
        // data, so that they hold no registers during the two stages
        float2 t_a[4][2], t_b[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const u32 m = (col * (4u * g + i)) & 65535u;
            t_a[i][0] = wn_hi[m >> 8];
            t_a[i][1] = wn_lo[m & 255u];
        }
        {
            const u32 m = (16u * col) & 65535u;
            t_b[0] = wn_hi[m >> 8];
            t_b[1] = wn_lo[m & 255u];
        }
        // (g,i) = k1, tile = k0 -> (g,i) = k0, tile = k1: four consecutive k
        // per lane, stored transposed as T[k + 256 c]
        float ta[16][4], tb[16][4];
        transpose_gi_tile(dr, ta);
        transpose_gi_tile(di, tb);
        cf fa[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const cf v = cmul(cf{t_a[i][0].x, sg * t_a[i][0].y}, cf{t_a[i][1].x, sg * t_a[i][1].y});
            fa[i] = cf{v.x * inv2, v.y * inv2};
        }
        const cf fstep = cmul(cf{t_b[0].x, sg * t_b[0].y}, cf{t_b[1].x, sg * t_b[1].y});
        cf fb = {1.f, 0.f};
        float2* q = dst + sbase + 256u * col + 4u * g;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            float o[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cf v = {ta[t][i], tb[t][i]};
                v = cmul(v, t ? cmul(fa[i], fb) : fa[i]);
                o[2 * i] = v.x;
                o[2 * i + 1] = v.y;
            }
            if (t < 15) fb = t ? cmul(fb, fstep) : fstep;
            const f32x4 o0 = {o[0], o[1], o[2], o[3]}, o1 = {o[4], o[5], o[6], o[7]};
            // cached stores: each instruction writes 16-byte pieces 32 bytes
            // apart (the next one fills the gaps), which L2 merges into whole
            // lines; non-temporal stores of the same pattern wrote 15% more
            // bytes and took 3.6% longer over the exec
            *reinterpret_cast<f32x4*>(q + 16u * t) = o0;
            *reinterpret_cast<f32x4*>(q + 16u * t + 2u) = o1;
        }
    }
}

template <int PASS, int SIGN>
__global__ __launch_bounds__(256, 2) void k_split_col(const float2* src, float2* dst,
                                                   const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                                   const float2* __restrict__ wn_lo, StartSched sched) {
    split_start_delay(sched);
    split_col_body<PASS, SIGN>(src, dst, blockIdx.x, w4096, wn_hi, wn_lo);
}

// Workgroups of 256 threads of kernel k resident at once on the current device.
template <class K>
static u32 split_resident(K k) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess)
        return 0;  // no schedule rather than a wrong one
    return (u32)(cus * per_cu);
}

template <int SIGN>
static hipError_t launch_split_t(const float2* in, float2* mid, float2* out, u32 n_signals, u32 chunk,
                                 const float2* w4096, const float2* wn_hi, const float2* wn_lo, StartSched sched,
                                 hipStream_t st) {
    // one cached count per kernel (each instantiation of this function and
    // each variable below is a distinct static)
    static const u32 res1 = split_resident(k_split_col<1, SIGN>);
    static const u32 res2 = split_resident(k_split_col<2, SIGN>);
    if (sched.groups == 0) sched.delay = 0;
    if (chunk == 0 || chunk > n_signals) chunk = n_signals;
    StartSched s1 = sched, s2 = sched;
    s1.resident = res1;
    s2.resident = res2;
    for (u32 s0 = 0; s0 < n_signals; s0 += chunk) {
        const u32 cnt = min(chunk, n_signals - s0);
        const size_t off = (size_t)s0 * 65536;
        hipLaunchKernelGGL((k_split_col<1, SIGN>), dim3(cnt * 4), dim3(256), 0, st, in + off, mid + off, w4096, wn_hi,
                           wn_lo, s1);
        hipLaunchKernelGGL((k_split_col<2, SIGN>), dim3(cnt * 4), dim3(256), 0, st, mid + off, out + off, w4096, wn_hi,
                           wn_lo, s2);
    }
    return hipGetLastError();
}

hipError_t launch_split_65536(const float2* in, float2* mid, float2* out, u32 n_signals, u32 chunk, int sign,
                              const float2* w4096, const float2* wn_hi, const float2* wn_lo, StartSched sched,
                              hipStream_t st) {
    return sign < 0 ? launch_split_t<-1>(in, mid, out, n_signals, chunk, w4096, wn_hi, wn_lo, sched, st)
                    : launch_split_t<+1>(in, mid, out, n_signals, chunk, w4096, wn_hi, wn_lo, sched, st);
}

// ---------------------------------------------------------------------------
// Every other shape: a sequence of column passes (SplitPassArgs).  A pass
// sees each signal as R rows x Q columns, element (r, q) at r*Q + q, and
// computes the R-point DFT down every column.  With q = q_hi*Qlo + q_lo it
// multiplies output k by W_M^{q_hi k} (M = R*Q/Qlo; no twiddle in the last
// pass of an axis) and stores it at q_hi*R*Qlo + k*Qlo + q_lo.  That is one
// step of the four-step recursion for the M-point DFTs of an [M][L] array
// along its first index (Qlo = L): x[c + (M/R) r][l] -> U[c][k][l] holds the
// (M/R)-point problems of U along c with R*L columns, so every pass again
// reads columns, and the last one (Q = Qlo) leaves natural order in place.
//
// A unit is 16 adjacent columns of one signal; a wave holds 4096 points:
// 256/R units of R rows (tile t = unit*R/16 + r0).  R = 16*F (F = 1..8):
//   stage 1 (tile r0, rows r = r0 + F r1, r1 = 4g + j): split DFT-16 over r1
//   -> rows k0 = 4g + i; twiddle W_R^{r0 k0}; DFT-F over the F tiles of the
//   unit in FP32 on VALU -> k = k0 + 16 k1 with k1 = the tile.
// R = 256: the two split MFMA stages of the 65536-point kernel above.
// Each unit takes its own power-of-two scale (wave_scale over its 16 columns),
// so units of different signals do not share one.

__device__ __forceinline__ cf split_tw(const SplitPassArgs& a, u32 e, float sg) {
    e &= a.m_mask;
    const float2 h = a.hi[e >> a.lo_bits], l = a.lo[e & ((1u << a.lo_bits) - 1u)];
    return cmul(cf{h.x, sg * h.y}, cf{l.x, sg * l.y});
}

__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <int R, bool ROW, int SIGN>
__device__ __forceinline__ void split_pass_f(const float2* src, float2* dst, const SplitPassArgs& a,
                                             const float2* __restrict__ w4096) {
    constexpr int F = R / 16, G = 16 / F, LR = F == 1 ? 4 : F == 2 ? 5 : F == 4 ? 6 : 7;
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id(), c16 = lane & 15, g = lane >> 4;
    const u32 wave = blockIdx.x * 4u + (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 upq = a.q_log2 - 4u;  // log2 of units per signal
    // the stage-1 twiddle steps W_R^{4g+i} first (global loads complete in
    // issue order: a table value loaded after the data would wait for it)
    float2 t_step[4];
    if constexpr (F > 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) t_step[i] = w4096[(4096 / R) * (4 * g + i)];
    }
    __builtin_amdgcn_sched_barrier(0);
    float dr[16][4], di[16][4];
#pragma unroll
    for (int cg = 0; cg < G; ++cg) {
        const u32 u = min(wave * G + cg, a.units - 1u);  // a ragged wave re-reads the last unit
        // wave-uniform base + 32-bit lane offset (< 2^27 elements)
        const float2* ub = src + ((size_t)(u >> upq) << (a.q_log2 + LR)) + ((u & ((1u << upq) - 1u)) << 4);
        const u32 lo = (u32)c16 + ((u32)(4 * F * g) << a.q_log2);
#pragma unroll
        for (int r0 = 0; r0 < F; ++r0)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float2* up = ub + ((size_t)(r0 + F * j) << a.q_log2);
                const f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(up + lo));
                dr[cg * F + r0][j] = v[0];
                di[cg * F + r0][j] = v[1];
            }
    }
    __builtin_amdgcn_sched_barrier(0);
    const Dft16Split Fop = make_dft16_split(lane, SIGN);
    float inv_u[G];
#pragma unroll
    for (int cg = 0; cg < G; ++cg) {
        float m = 0.f;
#pragma unroll
        for (int r0 = 0; r0 < F; ++r0)
#pragma unroll
            for (int j = 0; j < 4; ++j) m = fmaxf(m, fmaxf(fabsf(dr[cg * F + r0][j]), fabsf(di[cg * F + r0][j])));
        float s, inv;
        wave_scale(m, s, inv);
        s = uniform_f(s);
        inv_u[cg] = uniform_f(inv);
#pragma unroll
        for (int r0 = 0; r0 < F; ++r0) {
            const int t = cg * F + r0;
            f32x4 ore, oim;
            dft16_split_b<true>(dr[t], di[t], s, Fop, ore, oim);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dr[t][i] = ore[i];
                di[t][i] = oim[i];
            }
        }
    }
    if constexpr (F > 1) {
#pragma unroll
        for (int cg = 0; cg < G; ++cg)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf step = {t_step[i].x, sg * t_step[i].y};
                cf v[F];
                v[0] = cf{dr[cg * F][i], di[cg * F][i]};
                cf w = step;
#pragma unroll
                for (int r0 = 1; r0 < F; ++r0) {
                    v[r0] = cmul(cf{dr[cg * F + r0][i], di[cg * F + r0][i]}, w);
                    if (r0 + 1 < F) w = cmul(w, step);
                }
                small_dft<F>(v, SIGN);
#pragma unroll
                for (int k1 = 0; k1 < F; ++k1) {
                    dr[cg * F + k1][i] = v[k1].x;
                    di[cg * F + k1][i] = v[k1].y;
                }
            }
    }
    // outputs: tile cg*F + k1, word i -> k = 4g + i + 16 k1 of unit cg
#pragma unroll
    for (int cg = 0; cg < G; ++cg) {
        const u32 u = wave * G + cg;
        if (u >= a.units) break;  // wave-uniform
        const u32 q = ((u & ((1u << upq) - 1u)) << 4) + c16;
        float2* base = dst + ((size_t)(u >> upq) << (a.q_log2 + LR));
        const float inv = inv_u[cg];
        if constexpr (ROW) {
            // W_M^{q k}, M = R*Q, as W^{q(4g+i)} (W^{16q})^{k1}; row q of R
            // outputs, four consecutive k per lane (two 16-byte stores)
            cf fa[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf v = split_tw(a, q * (4u * g + i), sg);
                fa[i] = cf{v.x * inv, v.y * inv};
            }
            const cf fstep = split_tw(a, 16u * q, sg);
            cf fb = {1.f, 0.f};
            float2* o = base + ((size_t)q << LR) + 4 * g;
#pragma unroll
            for (int k1 = 0; k1 < F; ++k1) {
                float ov[8];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const cf v = cmul(cf{dr[cg * F + k1][i], di[cg * F + k1][i]}, k1 ? cmul(fa[i], fb) : fa[i]);
                    ov[2 * i] = v.x;
                    ov[2 * i + 1] = v.y;
                }
                if (k1 + 1 < F) fb = k1 ? cmul(fb, fstep) : fstep;
                const f32x4 o0 = {ov[0], ov[1], ov[2], ov[3]}, o1 = {ov[4], ov[5], ov[6], ov[7]};
                *reinterpret_cast<f32x4*>(o + 16 * k1) = o0;
                *reinterpret_cast<f32x4*>(o + 16 * k1 + 2) = o1;
            }
        } else {
            const u32 qhi = q >> a.qlo_log2, qlo = q & ((1u << a.qlo_log2) - 1u);
            float2* o = base + ((size_t)qhi << (a.qlo_log2 + LR)) + qlo;
            cf fa[4], fstep = {1.f, 0.f};
            const bool tw = a.m_mask != 0u;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf v = tw ? split_tw(a, qhi * (4u * g + i), sg) : cf{1.f, 0.f};
                fa[i] = cf{v.x * inv, v.y * inv};
            }
            if (tw) fstep = split_tw(a, 16u * qhi, sg);
            cf fb = {1.f, 0.f};
#pragma unroll
            for (int k1 = 0; k1 < F; ++k1) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const cf v = cmul(cf{dr[cg * F + k1][i], di[cg * F + k1][i]}, k1 ? cmul(fa[i], fb) : fa[i]);
                    const f32x2 w2 = {v.x, v.y};
                    __builtin_nontemporal_store(
                        w2, reinterpret_cast<f32x2*>(o + ((size_t)(4 * g + i + 16 * k1) << a.qlo_log2)));
                }
                if (k1 + 1 < F) fb = k1 ? cmul(fb, fstep) : fstep;
            }
        }
    }
}

template <bool ROW, int SIGN>
__device__ __forceinline__ void split_pass_256(const float2* src, float2* dst, const SplitPassArgs& a,
                                               const float2* __restrict__ w4096) {
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    const int lane = lane_id(), c16 = lane & 15, g = lane >> 4;
    const u32 wave = blockIdx.x * 4u + (u32)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const u32 upq = a.q_log2 - 4u;
    const u32 u = min(wave, a.units - 1u);
    const u32 q = ((u & ((1u << upq) - 1u)) << 4) + c16;
    const size_t sbase = (size_t)(u >> upq) << (a.q_log2 + 8);

    float2 t_step[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t_step[i] = w4096[16 * (4 * g + i)];
    __builtin_amdgcn_sched_barrier(0);
    // rows r = r0 + 16 (4g + j): tile r0, word j
    float dr[16][4], di[16][4];
    {
        // wave-uniform base + 32-bit lane offset (< 2^27 elements)
        const float2* ub = src + sbase + (q - c16);
        const u32 lo = (u32)c16 + ((u32)(64 * g) << a.q_log2);
#pragma unroll
        for (int r0 = 0; r0 < 16; ++r0)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float2* up = ub + ((size_t)(r0 + 16 * j) << a.q_log2);
                const f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(up + lo));
                dr[r0][j] = v[0];
                di[r0][j] = v[1];
            }
    }
    __builtin_amdgcn_sched_barrier(0);
    float m = 0.f;
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) m = fmaxf(m, fmaxf(fabsf(dr[t][j]), fabsf(di[t][j])));
    float s, inv;
    wave_scale(m, s, inv);
    s = uniform_f(s);
    inv = uniform_f(inv);
    {
        const Dft16Split F = make_dft16_split(lane, SIGN);
        cf step[4], rec[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            step[i] = cf{t_step[i].x, sg * t_step[i].y};
            rec[i] = cf{0.0625f, 0.f};
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 ore, oim;
            dft16_split_b<true>(dr[t], di[t], s, F, ore, oim);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf v = cmul(cf{ore[i], oim[i]}, rec[i]);
                if (t < 15) rec[i] = cmul(rec[i], step[i]);
                dr[t][i] = v.x;
                di[t][i] = v.y;
            }
        }
    }
    {
        float tr[16][4], ti[16][4];
        transpose_gi_tile(dr, tr);
        transpose_gi_tile(di, ti);
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dr[t][i] = tr[t][i];
                di[t][i] = ti[t][i];
            }
    }
    {
        // the DFT-16 operand again, from the opaque lane id (kept live from
        // stage 1 it would take registers through the transpose)
        const Dft16Split F2 = make_dft16_split(lane_id(), SIGN);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            f32x4 ore, oim;
            dft16_split_b<false>(dr[t], di[t], 1.f, F2, ore, oim);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dr[t][i] = ore[i];
                di[t][i] = oim[i];
            }
        }
    }
    if (wave >= a.units) return;  // wave-uniform; nothing to store
    // outputs: tile t = k0, (g, i) = k1 -> k = t + 16 (4g + i); inv2 undoes
    // the scale and the 1/16
    const float inv2 = inv * 16.f;
    if constexpr (ROW) {
        // -> (g, i) = k0, tile = k1: four consecutive k per lane, row q of 256
        float ta[16][4], tb[16][4];
        transpose_gi_tile(dr, ta);
        transpose_gi_tile(di, tb);
        cf fa[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const cf v = split_tw(a, q * (4u * g + i), sg);
            fa[i] = cf{v.x * inv2, v.y * inv2};
        }
        const cf fstep = split_tw(a, 16u * q, sg);
        cf fb = {1.f, 0.f};
        float2* o = dst + sbase + ((size_t)q << 8) + 4 * g;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            float ov[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf v = cmul(cf{ta[t][i], tb[t][i]}, t ? cmul(fa[i], fb) : fa[i]);
                ov[2 * i] = v.x;
                ov[2 * i + 1] = v.y;
            }
            if (t < 15) fb = t ? cmul(fb, fstep) : fstep;
            const f32x4 o0 = {ov[0], ov[1], ov[2], ov[3]}, o1 = {ov[4], ov[5], ov[6], ov[7]};
            *reinterpret_cast<f32x4*>(o + 16 * t) = o0;
            *reinterpret_cast<f32x4*>(o + 16 * t + 2) = o1;
        }
    } else {
        const u32 qhi = q >> a.qlo_log2, qlo = q & ((1u << a.qlo_log2) - 1u);
        float2* o = dst + sbase + ((size_t)qhi << (a.qlo_log2 + 8)) + qlo;
        // W_M^{q_hi k} = W^{16 q_hi (4g+i)} (W^{q_hi})^t
        cf fa[4], fstep = {1.f, 0.f};
        const bool tw = a.m_mask != 0u;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const cf v = tw ? split_tw(a, 16u * qhi * (4u * g + i), sg) : cf{1.f, 0.f};
            fa[i] = cf{v.x * inv2, v.y * inv2};
        }
        if (tw) fstep = split_tw(a, qhi, sg);
        cf fb = {1.f, 0.f};
#pragma unroll
        for (int t = 0; t < 16; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const cf v = cmul(cf{dr[t][i], di[t][i]}, t ? cmul(fa[i], fb) : fa[i]);
                const f32x2 w2 = {v.x, v.y};
                __builtin_nontemporal_store(w2,
                                            reinterpret_cast<f32x2*>(o + ((size_t)(t + 64 * g + 16 * i) << a.qlo_log2)));
            }
            if (t < 15) fb = t ? cmul(fb, fstep) : fstep;
        }
    }
}

template <int R, bool ROW, int SIGN>
__global__ __launch_bounds__(256, 2) void k_split_pass(const float2* src, float2* dst, SplitPassArgs a,
                                                    const float2* __restrict__ w4096, StartSched sched) {
    split_start_delay(sched);
    if constexpr (R == 256)
        split_pass_256<ROW, SIGN>(src, dst, a, w4096);
    else
        split_pass_f<R, ROW, SIGN>(src, dst, a, w4096);
}

// Radix R = 2, 4, 8 alone (the strided axis of a 2D shape with nx <= 8): one
// thread per column, FP32 on VALU (the only products are by +-1, +-i and
// (+-1 +-i)/sqrt(2)), no twiddle; in place or out of place.
template <int R, int SIGN>
__global__ __launch_bounds__(256) void k_split_small(const float2* src, float2* dst, u32 q_log2, u32 cols) {
    constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : 3;
    const u32 id = blockIdx.x * 256u + threadIdx.x;
    if (id >= cols) return;
    const size_t base = ((size_t)(id >> q_log2) << (q_log2 + LR)) + (id & ((1u << q_log2) - 1u));
    cf v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const f32x2 x = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(src + base + ((size_t)r << q_log2)));
        v[r] = cf{x[0], x[1]};
    }
    small_dft<R>(v, SIGN);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const f32x2 x = {v[k].x, v[k].y};
        __builtin_nontemporal_store(x, reinterpret_cast<f32x2*>(dst + base + ((size_t)k << q_log2)));
    }
}

template <int R, bool ROW, int SIGN>
static hipError_t launch_pass_t(const float2* src, float2* dst, const SplitPassArgs& a, const float2* w4096,
                                StartSched sched, hipStream_t st) {
    static const u32 res = split_resident(k_split_pass<R, ROW, SIGN>);
    constexpr u32 G = 256 / R;  // units per wave
    const u32 waves = (a.units + G - 1) / G;
    if (sched.groups == 0) sched.delay = 0;
    sched.resident = res;
    hipLaunchKernelGGL((k_split_pass<R, ROW, SIGN>), dim3((waves + 3) / 4), dim3(256), 0, st, src, dst, a, w4096,
                       sched);
    return hipGetLastError();
}

template <int R, int SIGN>
static hipError_t launch_small_t(const float2* src, float2* dst, const SplitPassArgs& a, hipStream_t st) {
    const u32 cols = a.units * 16u;
    hipLaunchKernelGGL((k_split_small<R, SIGN>), dim3((cols + 255) / 256), dim3(256), 0, st, src, dst, a.q_log2, cols);
    return hipGetLastError();
}

template <int SIGN>
static hipError_t launch_pass_s(int R, bool row, const float2* src, float2* dst, const SplitPassArgs& a,
                                const float2* w4096, StartSched sched, hipStream_t st) {
    switch (R) {
    case 2: return row ? hipErrorInvalidValue : launch_small_t<2, SIGN>(src, dst, a, st);
    case 4: return row ? hipErrorInvalidValue : launch_small_t<4, SIGN>(src, dst, a, st);
    case 8: return row ? hipErrorInvalidValue : launch_small_t<8, SIGN>(src, dst, a, st);
    case 16:
        return row ? launch_pass_t<16, true, SIGN>(src, dst, a, w4096, sched, st)
                   : launch_pass_t<16, false, SIGN>(src, dst, a, w4096, sched, st);
    case 32:
        return row ? launch_pass_t<32, true, SIGN>(src, dst, a, w4096, sched, st)
                   : launch_pass_t<32, false, SIGN>(src, dst, a, w4096, sched, st);
    case 64:
        return row ? launch_pass_t<64, true, SIGN>(src, dst, a, w4096, sched, st)
                   : launch_pass_t<64, false, SIGN>(src, dst, a, w4096, sched, st);
    case 128:
        return row ? launch_pass_t<128, true, SIGN>(src, dst, a, w4096, sched, st)
                   : launch_pass_t<128, false, SIGN>(src, dst, a, w4096, sched, st);
    case 256:
        return row ? launch_pass_t<256, true, SIGN>(src, dst, a, w4096, sched, st)
                   : launch_pass_t<256, false, SIGN>(src, dst, a, w4096, sched, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_split_pass(int R, bool row, const float2* src, float2* dst, const SplitPassArgs& a, int sign,
                             const float2* w4096, StartSched sched, hipStream_t st) {
    return sign < 0 ? launch_pass_s<-1>(R, row, src, dst, a, w4096, sched, st)
                    : launch_pass_s<+1>(R, row, src, dst, a, w4096, sched, st);
}

}  // namespace tcfft
