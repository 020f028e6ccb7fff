// gather.hip -- digit-reversal permutation on the device (natural-order input).
//
// Position p of a digit-reversed buffer holds input index rev(p), the mixed
// radix reversal that tcfft_doit_half.cpp:6-29 builds on the host: with p's
// digits p_0 (radix r_0, least significant) ... p_{m-1}, the index is
// p_{m-1} + r_{m-1}*(p_{m-2} + r_{m-2}*(... + r_1*p_0)).
#include "fft_kernels.hpp"

namespace tcfft {

__device__ __forceinline__ uint32_t rev_index(uint32_t p, const RevSpec& s) {
    uint32_t digits[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        if (i < s.n_digits) {
            digits[i] = p % (uint32_t)s.radix[i];
            p /= (uint32_t)s.radix[i];
        }
    }
    uint32_t idx = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i)
        if (i < s.n_digits) idx = idx * (uint32_t)s.radix[i] + digits[i];
    return idx;
}

__global__ __launch_bounds__(256) void k_gather_1d(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   uint32_t n_log2, uint32_t total, RevSpec spec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t mask = (1u << n_log2) - 1u;
    const uint32_t p = i & mask;
    out[i] = in[(i & ~mask) | rev_index(p, spec)];
}

__global__ __launch_bounds__(256) void k_gather_2d(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   uint32_t nx_log2, uint32_t ny_log2, uint32_t total, RevSpec sx,
                                                   RevSpec sy) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t y = i & ((1u << ny_log2) - 1u);
    const uint32_t x = (i >> ny_log2) & ((1u << nx_log2) - 1u);
    const uint32_t img = i >> (nx_log2 + ny_log2);
    const uint32_t src = (img << (nx_log2 + ny_log2)) | (rev_index(x, sx) << ny_log2) | rev_index(y, sy);
    out[i] = in[src];
}

hipError_t launch_gather_rev_1d(const uint32_t* in, uint32_t* out, uint32_t n_log2, uint32_t total,
                                const RevSpec& spec, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_1d, dim3((total + 255) / 256), dim3(256), 0, st, in, out, n_log2, total, spec);
    return hipGetLastError();
}

hipError_t launch_gather_rev_2d(const uint32_t* in, uint32_t* out, uint32_t nx_log2, uint32_t ny_log2,
                                uint32_t total, const RevSpec& sx, const RevSpec& sy, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_2d, dim3((total + 255) / 256), dim3(256), 0, st, in, out, nx_log2, ny_log2, total,
                       sx, sy);
    return hipGetLastError();
}

}  // namespace tcfft
