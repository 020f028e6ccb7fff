// conv.hip -- natural-order (decimation-in-frequency) transforms and the fused
// FFT convolution y = F^H (H * F x) / N for N = 4096*W (W = 1, 2, 4).
//
// The DIT kernels take digit-reversed input and produce natural output.  The
// DIF transform here is the other half: natural input, spectrum left in the
// same digit-reversed order that the DIT kernels consume, so a forward DIF,
// a pointwise product and an inverse DIT compose without any permutation.
// For N = 4096*W with radices [16, 16, 16, W] position p = d0 + 16*d1 +
// 256*d2 + 4096*d3 holds spectrum index k = d3 + W*(d2 + 16*d1 + 256*d0).
//
// DIF for N = 4096*W, n = n1 + 4096*n2, k = k2 + W*k1:
//   step A: Z_k2[n1] = W_N^{n1*k2} * sum_n2 x[n1 + 4096*n2] W_W^{n2*k2}
//           (VALU radix-W, 4 adjacent columns n1 per lane, coalesced loads),
//           then an LDS exchange so that wave k2 holds line Z_k2; the factor
//           W_N^{n1*k2} is applied inside the DIF's first stage (Dif4096Tw);
//   step B: Y[k2 + W*k1] = sum_n1 Z_k2[n1] W_4096^{n1*k1}, the in-wave
//           4096-point DIF below, whose output lands in registers exactly
//           where the DIT kernels load their input.
#include "fft_blocks.hpp"

namespace tcfft {

// ----------------------------------------------------------------------------
// In-wave 4096-point DIF, n1 = a + 16*b + 256*c, k1 = e + 16*f + 256*h:
//   X[k1] = sum_a W_16^{a*h} W_4096^{a*(e + 16f)} sum_b W_16^{b*f} W_256^{b*e}
//           sum_c W_16^{c*e} x[a + 16b + 256c]
// In : d[b][j] = tile b, lane (a = lane%16, g), word j: x[a + 16b + 256(4g + j)]
// Out: d[e][i] = tile e, lane%16 = f, (g,i) = h: X[e + 16f + 256h]
//   stage 1: B-form over c per tile b -> rows e; twiddle W_256^{b*e}
//            (recurrence over the tiles); transpose (g,i) <-> tile;
//   stage 2: A-form over b per tile e -> lane%16 = f, rows a; twiddle
//            W_4096^{a*(e + 16f)} = W_256^{a*f} * (W_4096^a)^e;
//   stage 3: B-form over a per tile e -> rows h; post(e, i, v) sees every
//            FP32 result before it is rounded.
// Line twiddle of the N = 4096*W transform (W > 1): line k2 of step A is
// Z_k2[n1] = W_N^{n1*k2} * (radix-W sum).  With n1 = a + 16b + 256c it
// splits into W_N^{256*k2*c} = W_{16W}^{k2*c}, a column scale of the stage-1
// operand (contraction index c = 4g + j), and W_N^{k2*a} * W_N^{16*k2*b},
// which does not depend on c and merges into the stage-1 recurrence: it
// starts at W_N^{k2*a} and steps by W_256^e * W_N^{16*k2} over the tiles b.
// So step A multiplies by nothing.
struct Dif4096Tw {
    float2 step1[4];  // W_256^{4g+i}
    float2 init2[4];  // W_256^{(4g+i)*f}, f = lane%16
    float2 step2[4];  // W_4096^{4g+i}
    float2 col[4];    // line twiddle: W_{16W}^{k2*(4g+j)}
    float2 s16;       // W_N^{16*k2} = W_{256W}^{k2}
    RawQ r0;          // W_N^{k2*a}, a = lane%16 (two-level table of W_N)
};

template <int W>
__device__ __forceinline__ void dif4096_tw_load(Dif4096Tw& tw, int lane, const float2* __restrict__ w4096, u32 k2,
                                                const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo,
                                                u32 wn_lo_bits) {
    const int c16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = 4 * g + i;
        tw.step1[i] = w4096[16 * a];
        tw.init2[i] = w4096[16 * ((a * c16) & 255)];
        tw.step2[i] = w4096[a];
    }
    if constexpr (W > 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) tw.col[j] = w4096[((k2 * (4u * g + j)) * (256u / W)) & 4095u];
        tw.s16 = w4096[(k2 * (16u / W)) & 4095u];
        const u32 m = (k2 * (u32)c16) & (4096u * W - 1u);
        tw.r0 = RawQ{wn_hi[m >> wn_lo_bits], wn_lo[m & ((1u << wn_lo_bits) - 1u)]};
    }
}

// DFT-16 operand whose contraction point n = 4g + j (this lane's g) is scaled
// by t[j]: D = F * diag(t) * X in the B-form.
__device__ __forceinline__ Dft16Op make_dft16_colscaled(int lane, int sign, const cf (&t)[4]) {
    const int f = lane & 15, g = lane >> 4;
    Dft16Op op;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = ((4 * g + j) * f) & 15;
        const float fr0 = cos16(m);
        const float fi0 = sign < 0 ? -sin16(m) : sin16(m);
        const float fr = fr0 * t[j].x - fi0 * t[j].y;
        const float fi = fr0 * t[j].y + fi0 * t[j].x;
        op.re[2 * j] = (_Float16)fr;
        op.re[2 * j + 1] = (_Float16)(-fi);
        op.im[2 * j] = (_Float16)fi;
        op.im[2 * j + 1] = (_Float16)fr;
    }
    return op;
}

template <int SIGN, int W, class Post>
__device__ __forceinline__ void dif4096(u32 (&d)[16][4], const Dif4096Tw& tw, const Dft16Op& op, Post post) {
    constexpr float sg = SIGN < 0 ? 1.f : -1.f;
    cf step[4], rec[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) step[i] = cf{tw.step1[i].x, sg * tw.step1[i].y};
    Dft16Op op1 = op;
    if constexpr (W > 1) {
        const cf s16 = {tw.s16.x, sg * tw.s16.y};
        const cf r0 = raw_q<SIGN>(tw.r0);
        cf t[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = cf{tw.col[j].x, sg * tw.col[j].y};
        op1 = make_dft16_colscaled(lane_id(), SIGN, t);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            step[i] = cmul(step[i], s16);
            rec[i] = r0;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) rec[i] = step[i];
    }
    // stage 1
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        f32x4 re, im;
        dft16_b(d[b], op1, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const cf v = {re[i], im[i]};
            if (W > 1 || b) {
                d[b][i] = pack(cmul(v, rec[i]));
                if (b < 15) rec[i] = cmul(rec[i], step[i]);
            } else {
                d[b][i] = pack(v);
            }
        }
    }
    u32 e[16][4];
    transpose_gi_tile(d, e);  // tile = e, lane%16 = a, (g,i) = b
    // stage 2
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        step[i] = cf{tw.step2[i].x, sg * tw.step2[i].y};
        rec[i] = cf{tw.init2[i].x, sg * tw.init2[i].y};
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_a(e[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            e[t][i] = pack(cmul(cf{re[i], im[i]}, rec[i]));
            if (t < 15) rec[i] = cmul(rec[i], step[i]);
        }
    }
    // stage 3
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        f32x4 re, im;
        dft16_b(e[t], op, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) d[t][i] = pack(post(t, i, cf{re[i], im[i]}));
    }
}

// ----------------------------------------------------------------------------
// Step A + LDS exchange.  Wave w of the signal's group handles columns
// n1 = 256*m + 4*lane + q, m = S*w + s (S = 16/W sweeps), for every line n2.
// Z_k2[n1] goes to LDS word 4096*k2 + zpos(n1) with n1 = a + 16b + 256c:
//   zpos = 4*((a ^ (b & 3)) + 16*b + 256*(c >> 2)) + ((c & 3) ^ ((b >> 2) & 3))
// so that a step-B reader (lane (a, g), tile b) finds its four words c = 4g + j
// in one 16-byte block (j permuted by a compile-time XOR), readers are
// conflict-free per 8 lanes and b32 writers spread over all banks.
__device__ __forceinline__ u32 dif_zpos(u32 n1) {
    const u32 a = n1 & 15u, b = (n1 >> 4) & 15u, c = n1 >> 8;
    return 4u * ((a ^ (b & 3u)) + 16u * b + 256u * (c >> 2)) + ((c & 3u) ^ ((b >> 2) & 3u));
}

// Loads of the natural-order signal at src for step A (16 x 16 B per lane).
template <int W>
__device__ __forceinline__ void stepa_load(u32 (&v)[16][4], const u32* __restrict__ src, int w, int lane) {
    constexpr int S = 16 / W;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int n2 = 0; n2 < W; ++n2) load4_nt(v[s * W + n2], src + 256u * (S * w + s) + 4u * lane + 4096u * n2);
}

// Radix-W over the W lines for this wave's columns (no twiddle: the line
// twiddle is folded into the DIF, see Dif4096Tw), written to LDS line k2.
template <int W, int SIGN>
__device__ __forceinline__ void stepa_exchange(u32 (&v)[16][4], int w, int lane, u32* lds) {
    constexpr int S = 16 / W;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const u32 m = (u32)(S * w + s);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u32 zp = dif_zpos(256u * m + 4u * lane + q);
            if constexpr (W == 2) {
                // radix 2 on the packed words (packed fp16 add and subtract)
                lds[zp] = hadd2(v[s * W][q], v[s * W + 1][q]);
                lds[4096u + zp] = hsub2(v[s * W][q], v[s * W + 1][q]);
            } else {
                cf x[W];
#pragma unroll
                for (int n2 = 0; n2 < W; ++n2) x[n2] = unpack(v[s * W + n2][q]);
                if constexpr (W > 1) small_dft<W>(x, SIGN);
#pragma unroll
                for (int k = 0; k < W; ++k) lds[4096u * k + zp] = pack(x[k]);
            }
        }
    }
}

// Step-B read of line k2 = w into the dif4096 input layout.
__device__ __forceinline__ void stepb_read(u32 (&d)[16][4], const u32* lds, int w, int lane) {
    const u32 a = lane & 15, g = lane >> 4;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        u32 q[4];
        load4(q, lds + 4096u * w + 4u * ((a ^ (b & 3u)) + 16u * b + 256u * g));
#pragma unroll
        for (int j = 0; j < 4; ++j) d[b][j] = q[j ^ ((b >> 2) & 3)];
    }
}

// ----------------------------------------------------------------------------
// k_dif: natural-order x (N = 4096*W per signal) -> spectrum in the drop-in's
// digit-reversed order, out of place (in != out) or in place.
template <int W, int SIGN>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_dif(const u32* in, u32* out,
                                                const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                                const float2* __restrict__ wn_lo, u32 wn_lo_bits) {
    extern __shared__ u32 lds[];
    constexpr u32 N = 4096u * W;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int c16 = lane & 15, g = lane >> 4;
    Dif4096Tw dtw;
    dif4096_tw_load<W>(dtw, lane, w4096, (u32)w, wn_hi, wn_lo, wn_lo_bits);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
    stepa_load<W>(d, in + (size_t)blockIdx.x * N, w, lane);
    __builtin_amdgcn_sched_barrier(0);
    stepa_exchange<W, SIGN>(d, w, lane, lds);
    __syncthreads();
    stepb_read(d, lds, w, lane);
    const Dft16Op op = make_dft16(lane, SIGN);
    dif4096<SIGN, W>(d, dtw, op, [](int, int, cf v) -> cf { return v; });
    u32* dst = out + (size_t)blockIdx.x * N + 4096u * w + 16u * c16 + 4u * g;
#pragma unroll
    for (int t = 0; t < 16; ++t) store4_nt(dst + 256 * t, d[t]);
}

// ----------------------------------------------------------------------------
// k_conv: y = F^H (H * F x) * scale for one signal per workgroup of W waves.
// x, y: [batch][channels][N] fp16 complex, natural order; h_rev: [channels][N]
// fp16 complex in the digit-reversed spectrum order (conjugated with CONJ);
// with SPEC the spectrum F x is also stored, digit-reversed, at spec.
// The product is taken on the FP32 results of the last DIF stage; the
// inverse is the DIT of the single-pass kernel (sign +1), with the scale
// carried by its stage-1 twiddle recurrence (FP32, before rounding).
// x and y may alias (in place), so neither is __restrict__: a workgroup reads
// its signal completely before its first store of y.
template <int W, bool CONJ, bool SPEC>
__global__ __launch_bounds__(64 * W) TCFFT_MFMA_BUDGET void k_conv(const u32* x, const u32* __restrict__ h_rev,
                                                 u32* y, u32* __restrict__ spec, u32 channels,
                                                 float scale, const float2* __restrict__ w4096,
                                                 const float2* __restrict__ wn_hi, const float2* __restrict__ wn_lo,
                                                 u32 wn_lo_bits) {
    extern __shared__ u32 lds[];
    constexpr u32 N = 4096u * W;
    const u32 sig = blockIdx.x;
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const int c16 = lane & 15, g = lane >> 4;
    Dif4096Tw dtw;
    dif4096_tw_load<W>(dtw, lane, w4096, (u32)w, wn_hi, wn_lo, wn_lo_bits);
    __builtin_amdgcn_sched_barrier(0);
    u32 d[16][4];
    stepa_load<W>(d, x + (size_t)sig * N, w, lane);
    __builtin_amdgcn_sched_barrier(0);
    stepa_exchange<W, -1>(d, w, lane, lds);
    __syncthreads();
    stepb_read(d, lds, w, lane);

    // filter tiles of this line, loaded four tiles ahead of their use
    const u32 lane_off = 4096u * w + 16u * c16 + 4u * g;
    const u32* hp = h_rev + (size_t)(sig % channels) * N + lane_off;
    u32* sp = SPEC ? spec + (size_t)sig * N + lane_off : nullptr;
    u32 hq[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) load4(hq[t], hp + 256 * t);
    {
        const Dft16Op op = make_dft16(lane, -1);
        u32 sq[4];
        dif4096<-1, W>(d, dtw, op, [&](int t, int i, cf v) -> cf {
            const cf h = unpack(hq[t & 3][i]);
            if (i == 3 && t + 4 < 16) load4(hq[t & 3], hp + 256 * (t + 4));
            if constexpr (SPEC) {
                sq[i] = pack(v);
                if (i == 3) store4_nt(sp + 256 * t, sq);
            }
            return CONJ ? cf{v.x * h.x + v.y * h.y, v.y * h.x - v.x * h.y} : cmul(v, h);
        });
    }

    // inverse (sign +1): the single-pass DIT of the line, exchange, radix-W stage
    BigTw bt;
    big_tw_load<W, false>(bt, lane, w, sig, w4096, wn_hi, wn_lo, wn_lo_bits, EpiParams{});
    const Dft16Op opi = make_dft16(lane, +1);
    u32 e[16][4];
    big_phase1<+1>(d, e, bt, opi, scale);  // the 1/n scale rides on the stage-1 recurrence
    __syncthreads();  // every wave has read its step-B line before the exchange reuses the LDS
    big_phase2<W, 1, false, +1>(e, w, lds, y + (size_t)sig * N, bt, opi, true);
}

// ----------------------------------------------------------------------------
// Filter gradient: gh_rev[c][p] = sum_b conj(X[b][c][p]) * G[b][c][p] * scale,
// X, G digit-reversed fp16 spectra [batch][channels][N], gh_rev FP32.
__global__ __launch_bounds__(256) void k_conv_wgrad(const u32* __restrict__ xs, const u32* __restrict__ gs,
                                                    float2* __restrict__ gh, u32 n_per_channel, u32 channels,
                                                    u32 batch, float scale) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;  // over channels * N / 2 pairs
    const size_t total = (size_t)channels * n_per_channel;
    const size_t p0 = idx * 2;
    if (p0 >= total) return;
    cf acc0 = {0.f, 0.f}, acc1 = {0.f, 0.f};
    for (u32 b = 0; b < batch; ++b) {
        const size_t off = (size_t)b * total + p0;
        const uint2 xv = *reinterpret_cast<const uint2*>(xs + off);
        const uint2 gv = *reinterpret_cast<const uint2*>(gs + off);
        const cf x0 = unpack(xv.x), x1 = unpack(xv.y), g0 = unpack(gv.x), g1 = unpack(gv.y);
        acc0.x += x0.x * g0.x + x0.y * g0.y;
        acc0.y += x0.x * g0.y - x0.y * g0.x;
        acc1.x += x1.x * g1.x + x1.y * g1.y;
        acc1.y += x1.x * g1.y - x1.y * g1.x;
    }
    gh[p0] = make_float2(acc0.x * scale, acc0.y * scale);
    gh[p0 + 1] = make_float2(acc1.x * scale, acc1.y * scale);
}

// ----------------------------------------------------------------------------
template <class K>
static hipError_t conv_set_lds(K kernel, size_t bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int W, int SIGN>
static hipError_t launch_dif_t(const u32* in, u32* out, u32 n_signals, const float2* w4096, const float2* wn_hi,
                               const float2* wn_lo, u32 lo_bits, hipStream_t st) {
    const size_t lds = 4096u * W * sizeof(u32);
    auto k = k_dif<W, SIGN>;
    static hipError_t attr = conv_set_lds(k, lds);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, in, out, w4096, wn_hi, wn_lo, lo_bits);
    return hipGetLastError();
}

hipError_t launch_dif(int W, const uint32_t* in, uint32_t* out, uint32_t n_signals, int sign, const float2* w4096,
                      const float2* wn_hi, const float2* wn_lo, uint32_t lo_bits, hipStream_t st) {
    switch (W * 4 + (sign < 0 ? 0 : 1)) {
    case 4: return launch_dif_t<1, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    case 5: return launch_dif_t<1, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    case 8: return launch_dif_t<2, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    case 9: return launch_dif_t<2, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    case 16: return launch_dif_t<4, -1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    case 17: return launch_dif_t<4, +1>(in, out, n_signals, w4096, wn_hi, wn_lo, lo_bits, st);
    default: return hipErrorInvalidValue;
    }
}

template <int W, bool CONJ, bool SPEC>
static hipError_t launch_conv_t(const u32* x, const u32* h_rev, u32* y, u32* spec, u32 n_signals, u32 channels,
                                float scale, const float2* w4096, const float2* wn_hi, const float2* wn_lo,
                                u32 lo_bits, hipStream_t st) {
    const size_t lds = 4096u * W * sizeof(u32);
    auto k = k_conv<W, CONJ, SPEC>;
    static hipError_t attr = conv_set_lds(k, lds);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(k, dim3(n_signals), dim3(64 * W), lds, st, x, h_rev, y, spec, channels, scale, w4096, wn_hi,
                       wn_lo, lo_bits);
    return hipGetLastError();
}

template <int W>
static hipError_t launch_conv_w(bool conj, const u32* x, const u32* h_rev, u32* y, u32* spec, u32 n_signals,
                                u32 channels, float scale, const float2* w4096, const float2* wn_hi,
                                const float2* wn_lo, u32 lo_bits, hipStream_t st) {
    if (conj)
        return spec ? launch_conv_t<W, true, true>(x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo,
                                                   lo_bits, st)
                    : launch_conv_t<W, true, false>(x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi,
                                                    wn_lo, lo_bits, st);
    return spec ? launch_conv_t<W, false, true>(x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo,
                                                lo_bits, st)
                : launch_conv_t<W, false, false>(x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo,
                                                 lo_bits, st);
}

hipError_t launch_conv(int W, bool conj, const uint32_t* x, const uint32_t* h_rev, uint32_t* y, uint32_t* spec,
                       uint32_t n_signals, uint32_t channels, float scale, const float2* w4096, const float2* wn_hi,
                       const float2* wn_lo, uint32_t lo_bits, hipStream_t st) {
    switch (W) {
    case 1: return launch_conv_w<1>(conj, x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo, lo_bits, st);
    case 2: return launch_conv_w<2>(conj, x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo, lo_bits, st);
    case 4: return launch_conv_w<4>(conj, x, h_rev, y, spec, n_signals, channels, scale, w4096, wn_hi, wn_lo, lo_bits, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_conv_wgrad(const uint32_t* xs, const uint32_t* gs, float2* gh, uint32_t n, uint32_t channels,
                             uint32_t batch, float scale, hipStream_t st) {
    const size_t pairs = (size_t)channels * n / 2;
    const u32 blocks = (u32)((pairs + 255) / 256);
    hipLaunchKernelGGL(k_conv_wgrad, dim3(blocks), dim3(256), 0, st, xs, gs, gh, n, channels, batch, scale);
    return hipGetLastError();
}

}  // namespace tcfft
