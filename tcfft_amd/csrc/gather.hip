// gather.hip -- digit-reversal permutation on the device (natural-order input).
//
// Position p of a digit-reversed buffer holds input index rev(p), the mixed
// radix reversal that tcfft_doit_half.cpp:6-29 builds on the host: with p's
// digits p_0 (radix r_0, least significant) ... p_{m-1}, the index is
// p_{m-1} + r_{m-1}*(p_{m-2} + r_{m-2}*(... + r_1*p_0)).
#include "fft_kernels.hpp"

namespace tcfft {

// Radices are powers of two: digits by shifts and masks (an integer division
// by a runtime value costs dozens of instructions per digit).
__device__ __forceinline__ uint32_t rev_index(uint32_t p, const RevSpec& s) {
    uint32_t digits[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        if (i < s.n_digits) {
            const uint32_t w = (uint32_t)__builtin_ctz((uint32_t)s.radix[i]);
            digits[i] = p & ((1u << w) - 1u);
            p >>= w;
        }
    }
    uint32_t idx = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i)
        if (i < s.n_digits) idx = (idx << __builtin_ctz((uint32_t)s.radix[i])) | digits[i];
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

// Tiled 1D digit reversal through LDS.  Position p's bit permutation: bit i
// of rev(p) is bit pi(i) of p.  A tile of 2^(a+b) elements fixes every bit
// of p except the low a bits (consecutive positions: the writes) and the b
// bits pi(0..b-1) (consecutive input indices: the reads), which the host
// chose disjoint.  Tile element tau: its low a bits are p's low a bits, its
// high b bits are rev(p)'s low b bits.  LDS word tau ^ ((tau >> a) & (2^a-1))
// keeps both the column-wise writes and the row-wise reads conflict-free.
__global__ __launch_bounds__(256) void k_gather_tiled(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint32_t n_log2, RevSpec spec, GatherTile gt) {
    __shared__ uint32_t tile[4096];
    const uint32_t elems = 1u << (gt.a + gt.b);
    const uint32_t tiles_per_signal = 1u << (n_log2 - gt.a - gt.b);
    const uint32_t sig = blockIdx.x / tiles_per_signal;
    uint32_t f = blockIdx.x % tiles_per_signal;
    // deposit the tile index into the free bits of p
    uint32_t p_block = 0, mask = gt.free_mask;
    while (mask) {
        const uint32_t bit = (uint32_t)__builtin_ctz(mask);
        p_block |= (f & 1u) << bit;
        f >>= 1;
        mask &= mask - 1u;
    }
    const size_t base = (size_t)sig << n_log2;
    const uint32_t amask = (1u << gt.a) - 1u, bmask = (1u << gt.b) - 1u;
    // rev is a bit permutation: rev(p) = rev(p_block) | (tau >> a) | the low
    // a bits of p moved to their places in rev (low_to_rev)
    const uint32_t rev_block = rev_index(p_block, spec);
    auto pos_of = [&](uint32_t tau) -> uint32_t {
        uint32_t p = p_block | (tau & amask);
        const uint32_t hi = tau >> gt.a;
#pragma unroll
        for (uint32_t i = 0; i < 6; ++i)
            if (i < gt.b) p |= ((hi >> i) & 1u) << gt.in_bit[i];
        return p;
    };
    auto rev_of = [&](uint32_t tau) -> uint32_t {
        uint32_t r = rev_block | (tau >> gt.a);
#pragma unroll
        for (uint32_t j = 0; j < 6; ++j)
            if (j < gt.a) r |= ((tau >> j) & 1u) << gt.low_to_rev[j];
        return r;
    };
    auto lds_word = [&](uint32_t tau) { return tau ^ ((tau >> gt.a) & amask); };
    for (uint32_t e = threadIdx.x; e < elems; e += 256u) {
        const uint32_t tau = ((e & bmask) << gt.a) | (e >> gt.b);  // consecutive e: consecutive rev(p)
        tile[lds_word(tau)] = in[base + rev_of(tau)];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < elems; e += 256u) out[base + pos_of(e)] = tile[lds_word(e)];
}

hipError_t launch_gather_tiled_1d(const uint32_t* in, uint32_t* out, uint32_t n_log2, uint32_t total,
                                  const RevSpec& spec, const GatherTile& gt, hipStream_t st) {
    if (gt.a + gt.b > 12 || gt.a + gt.b > n_log2 || gt.b > 6) return hipErrorInvalidValue;
    const uint32_t blocks = total >> (gt.a + gt.b);
    hipLaunchKernelGGL(k_gather_tiled, dim3(blocks), dim3(256), 0, st, in, out, n_log2, spec, gt);
    return hipGetLastError();
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
