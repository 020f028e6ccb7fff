// split.hip -- split-precision FFT: FP32 complex in/out, FP16 hi+lo on MFMA.
//
// Every radix-16 stage computes D = Fh*Xh + Fh*Xl + Fl*Xh with
// v_mfma_f32_16x16x32_f16 and FP32 accumulation, where X = Xh + Xl is the
// FP16 hi/lo split of the (scaled) FP32 data and F = Fh + Fl that of the
// DFT-16 matrix.  The dropped Fl*Xl term and the two FP16 roundings leave
// about 2^-22 relative error per stage.  Before each split the 16 points of
// an MFMA column are scaled by a power of two so their largest component is
// below 2^10 (no subnormal lo parts, no overflow); the scale is undone on the
// FP32 result, exactly.  Twiddles come from FP64-generated tables.
//
// N = 65536 = 256 x 256 as a natural-order four-step in two passes over
// FP32 data, each pass a workgroup-per-(signal, 16 columns) kernel staging a
// 256 x 16 tile in LDS:
//   pass 1: x[c + 256 r] -> Y_c[k] = sum_r x W_256^{rk}, times W_N^{c k},
//           stored transposed at out[k + 256 c]
//   pass 2: out[k + 256 c] -> X[k + 256 k'] = sum_c Y'_c[k] W_256^{c k'},
//           in place
// Inside a pass the 256-point DFT is 16 x 16:
//   X[k0 + 16 k1] = sum_r0 W_16^{r0 k1} W_256^{r0 k0} sum_r1 x[r0 + 16 r1] W_16^{r1 k0}.
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

// DFT-16 (B-form) of the lane's four points v, scaled per MFMA column.
// Returns the FP32 results for frequencies 4g+i of column lane%16.
__device__ __forceinline__ void dft16_split(const cf* v, const Dft16Split& F, cf* out) {
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) m = fmaxf(m, fmaxf(fabsf(v[j].x), fabsf(v[j].y)));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    int e = 0;
    (void)frexpf(m, &e);  // m < 2^e
    const float s = m > 0.f ? ldexpf(1.f, 10 - e) : 1.f;
    const float inv = m > 0.f ? ldexpf(1.f, e - 10) : 1.f;
    u32 hi[4], lo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float xr = v[j].x * s, xi = v[j].y * s;
        hi[j] = pack(xr, xi);
        const cf h = unpack(hi[j]);
        lo[j] = pack(xr - h.x, xi - h.y);
    }
    const f16x8 bh = as_operand(hi[0], hi[1], hi[2], hi[3]);
    const f16x8 bl = as_operand(lo[0], lo[1], lo[2], lo[3]);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    f32x4 re = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.re, bh, z, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.re, bl, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.l.re, bh, re, 0, 0, 0);
    f32x4 im = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.im, bh, z, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.h.im, bl, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.l.im, bh, im, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = cf{re[i] * inv, im[i] * inv};
}

constexpr int SPLIT_LDS_STRIDE = 17;  // float2 per LDS row (16 columns + 1 pad)

template <int MODE, int SIGN>
__global__ __launch_bounds__(256) void k_split256(const float2* __restrict__ src, float2* __restrict__ dst,
                                                  const float2* __restrict__ w4096, const float2* __restrict__ wn_hi,
                                                  const float2* __restrict__ wn_lo) {
    constexpr int sign = SIGN;
    constexpr u32 N = 65536;
    __shared__ float2 tile[256 * SPLIT_LDS_STRIDE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int c16 = lane & 15, g = lane >> 4;
    const float sg = sign < 0 ? 1.f : -1.f;
    const u32 b = blockIdx.x >> 4, q = blockIdx.x & 15;
    const size_t sbase = (size_t)b * N;

    // load 256 rows x 16 columns (columns 16q..16q+15, row stride 256)
    {
        const float2* in = (MODE == 1 ? src : dst) + sbase + 16u * q;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int row = (tid >> 3) + 32 * j, cp = tid & 7;
            const f32x4* p = reinterpret_cast<const f32x4*>(in + (size_t)row * 256 + 2 * cp);
            const f32x4 v = MODE == 1 ? __builtin_nontemporal_load(p) : *p;
            tile[row * SPLIT_LDS_STRIDE + 2 * cp] = make_float2(v[0], v[1]);
            tile[row * SPLIT_LDS_STRIDE + 2 * cp + 1] = make_float2(v[2], v[3]);
        }
    }
    __syncthreads();
    const Dft16Split F = make_dft16_split(lane, sign);

    // stage A: tile r0 = 4w+u, DFT over r1 (rows r0 + 16 r1), twiddle W_256^{r0 k0}
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int r0 = 4 * w + u;
        cf v[4], o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float2 t = tile[(r0 + 16 * (4 * g + i)) * SPLIT_LDS_STRIDE + c16];
            v[i] = cf{t.x, t.y};
        }
        dft16_split(v, F, o);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k0 = 4 * g + i;
            const cf x = cmul(o[i], twiddle(w4096, ((r0 * k0) & 255) * 16, sg));
            tile[(r0 + 16 * k0) * SPLIT_LDS_STRIDE + c16] = make_float2(x.x, x.y);
        }
    }
    __syncthreads();
    // stage B: tile k0 = 4w+u, DFT over r0 (rows 16 k0 + r0) -> X[k0 + 16 k1]
    cf res[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int k0 = 4 * w + u;
        cf v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float2 t = tile[(16 * k0 + 4 * g + i) * SPLIT_LDS_STRIDE + c16];
            v[i] = cf{t.x, t.y};
        }
        dft16_split(v, F, res[u]);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const u32 k0 = 4u * w + u;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const u32 k = k0 + 16u * (4u * g + i);
            cf x = res[u][i];
            if constexpr (MODE == 1) {
                const u32 c = 16u * q + c16;
                const u32 m = (c * k) & (N - 1u);
                x = cmul(x, cmul(twiddle(wn_hi, m >> 8, sg), twiddle(wn_lo, m & 255u, sg)));
            }
            tile[k * SPLIT_LDS_STRIDE + c16] = make_float2(x.x, x.y);
        }
    }
    __syncthreads();
    if constexpr (MODE == 1) {
        // out row c = 16q + cl holds Y'_c[k], k = 0..255 contiguous
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int idx = tid + 256 * j;
            const int cl = idx >> 7, kp = idx & 127;
            const float2 a = tile[(2 * kp) * SPLIT_LDS_STRIDE + cl];
            const float2 bb = tile[(2 * kp + 1) * SPLIT_LDS_STRIDE + cl];
            f32x4* p = reinterpret_cast<f32x4*>(dst + sbase + (size_t)(16u * q + cl) * 256 + 2 * kp);
            const f32x4 v = {a.x, a.y, bb.x, bb.y};
            __builtin_nontemporal_store(v, p);
        }
    } else {
        float2* out = dst + sbase + 16u * q;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int row = (tid >> 3) + 32 * j, cp = tid & 7;
            const float2 a = tile[row * SPLIT_LDS_STRIDE + 2 * cp];
            const float2 bb = tile[row * SPLIT_LDS_STRIDE + 2 * cp + 1];
            const f32x4 v = {a.x, a.y, bb.x, bb.y};
            __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(out + (size_t)row * 256 + 2 * cp));
        }
    }
}

template <int SIGN>
static hipError_t launch_split_t(const float2* in, float2* out, u32 n_signals, u32 chunk, const float2* w4096,
                                 const float2* wn_hi, const float2* wn_lo, hipStream_t st) {
    if (chunk == 0 || chunk > n_signals) chunk = n_signals;
    for (u32 s0 = 0; s0 < n_signals; s0 += chunk) {
        const u32 cnt = min(chunk, n_signals - s0);
        const size_t off = (size_t)s0 * 65536;
        hipLaunchKernelGGL((k_split256<1, SIGN>), dim3(cnt * 16), dim3(256), 0, st, in + off, out + off, w4096, wn_hi,
                           wn_lo);
        hipLaunchKernelGGL((k_split256<2, SIGN>), dim3(cnt * 16), dim3(256), 0, st, out + off, out + off, w4096, wn_hi,
                           wn_lo);
    }
    return hipGetLastError();
}

hipError_t launch_split_65536(const float2* in, float2* out, u32 n_signals, u32 chunk, int sign, const float2* w4096,
                              const float2* wn_hi, const float2* wn_lo, hipStream_t st) {
    return sign < 0 ? launch_split_t<-1>(in, out, n_signals, chunk, w4096, wn_hi, wn_lo, st)
                    : launch_split_t<+1>(in, out, n_signals, chunk, w4096, wn_hi, wn_lo, st);
}

}  // namespace tcfft
