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

// Tiled 1D digit reversal through LDS.  Position p's bit permutation: bit i
// of rev(p) is bit pi(i) of p.  A tile of 2^(a+b) elements fixes every bit
// of p except the low a bits (consecutive positions: the writes) and the b
// bits pi(0..b-1) (consecutive input indices: the reads), which the host
// chose disjoint (a >= 2, b >= 2).  Tile element tau: its low a bits are p's
// low a bits, its high b bits are rev(p)'s low b bits, so
//   p(tau)   = p_block   | (tau & (2^a-1)) | pos_hi[tau >> a]
//   rev(tau) = rev_block | (tau >> a)      | rev_lo[tau & (2^a-1)]
// with the two 64-entry deposit tables built once per workgroup in LDS.
// Each thread moves 4 consecutive elements with one 16-byte load (4
// consecutive rev) and one 16-byte store (4 consecutive p).  LDS word
// tau ^ ((tau >> a) & (2^a-1)) spreads both phases over the banks.
__global__ __launch_bounds__(256) void k_gather_tiled(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint32_t n_log2, RevSpec spec, GatherTile gt) {
    __shared__ uint32_t tile[4096];
    __shared__ uint32_t pos_hi[64], rev_lo[64];
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
    const uint32_t rev_block = rev_index(p_block, spec);
    const uint32_t amask = (1u << gt.a) - 1u, bmask = (1u << gt.b) - 1u;
    {
        const uint32_t t = threadIdx.x & 63u;
        uint32_t v = 0;
        if (threadIdx.x < 64u) {
            for (uint32_t i = 0; i < gt.b; ++i) v |= ((t >> i) & 1u) << gt.in_bit[i];
            pos_hi[t] = v;
        } else if (threadIdx.x < 128u) {
            for (uint32_t j = 0; j < gt.a; ++j) v |= ((t >> j) & 1u) << gt.low_to_rev[j];
            rev_lo[t] = v;
        }
    }
    __syncthreads();
    const size_t base = (size_t)sig << n_log2;
    auto lds_word = [&](uint32_t tau) { return tau ^ ((tau >> gt.a) & amask); };
    for (uint32_t e = 4u * threadIdx.x; e < elems; e += 1024u) {
        // e .. e+3: consecutive input indices (the low b bits of rev)
        const uint32_t hi = e >> gt.b;  // tau's low a bits
        const uint32_t src = rev_block | (e & bmask) | rev_lo[hi & amask];
        const uint4 v = *reinterpret_cast<const uint4*>(in + base + src);
        const uint32_t t0 = ((e & bmask) << gt.a) | hi;
        tile[lds_word(t0)] = v.x;
        tile[lds_word(t0 + (1u << gt.a))] = v.y;
        tile[lds_word(t0 + (2u << gt.a))] = v.z;
        tile[lds_word(t0 + (3u << gt.a))] = v.w;
    }
    __syncthreads();
    for (uint32_t e = 4u * threadIdx.x; e < elems; e += 1024u) {
        // tau = e .. e+3: consecutive positions (the low a bits of p)
        const uint32_t dst = p_block | (e & amask) | pos_hi[e >> gt.a];
        uint4 v;
        v.x = tile[lds_word(e)];
        v.y = tile[lds_word(e + 1u)];
        v.z = tile[lds_word(e + 2u)];
        v.w = tile[lds_word(e + 3u)];
        *reinterpret_cast<uint4*>(out + base + dst) = v;
    }
}

hipError_t launch_gather_tiled_1d(const uint32_t* in, uint32_t* out, uint32_t n_log2, uint32_t total,
                                  const RevSpec& spec, const GatherTile& gt, hipStream_t st) {
    if (gt.a + gt.b > 12 || gt.a + gt.b > n_log2 || gt.a > 6 || gt.b > 6 || gt.a < 2 || gt.b < 2)
        return hipErrorInvalidValue;
    const uint32_t blocks = total >> (gt.a + gt.b);
    hipLaunchKernelGGL(k_gather_tiled, dim3(blocks), dim3(256), 0, st, in, out, n_log2, spec, gt);
    return hipGetLastError();
}

}  // namespace tcfft
