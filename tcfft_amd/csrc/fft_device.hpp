// fft_device.hpp -- wave64 building blocks for the gfx950 FP16 FFT kernels.
//
// Storage type: one complex FP16 element = one 32-bit word (re in bits 0-15,
// im in bits 16-31), the interleaved `half` pairs of tcfft_half.h's buffers.
//
// The radix-16 DFT runs on v_mfma_f32_16x16x32_f16 with the complex product
// written as a real block product.  Operand layouts (wave64, verified by a
// known-answer probe on MI355X):
//   A (16x32): lane l holds row m = l%16, k = 8*(l/16) .. 8*(l/16)+7
//   B (32x16): lane l holds col n = l%16, k = 8*(l/16) .. 8*(l/16)+7
//   D (16x16): lane l holds col n = l%16, rows m = 4*(l/16) + i, i = 0..3
// The data operand is four consecutive complex words of a lane: K-slot
// 8g+2j+c is component c (0 re, 1 im) of point 4g+j, g = l/16.  The constant
// operand pairs (re_op, im_op) hold, for point n = 4g+j and frequency f = l%16,
//   re_op: (Fr, -Fi)   im_op: (Fi, Fr),   F(n,f) = exp(sign*2*pi*i*n*f/16)
// so Out_re = sum xr*Fr - xi*Fi and Out_im = sum xr*Fi + xi*Fr.  The same
// registers serve both operand positions because F is symmetric:
//   A-form (data as A): out lane%16 = frequency, rows 4g+i = data column
//   B-form (data as B): out lane%16 = data column, rows 4g+i = frequency
// This replaces the 4-WMMA + negate complex_mul of tcfft_half.cu:5-20 with two
// MFMAs that accumulate in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcfft {

typedef uint32_t u32;
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef u32 u32x4 __attribute__((ext_vector_type(4)));
typedef u32 u32x2 __attribute__((ext_vector_type(2)));

struct cf {  // FP32 complex value in registers
    float x, y;
};

__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }

// Round-to-nearest FP32 pair -> packed FP16 word (one v_cvt_pk_f16_f32).
__device__ __forceinline__ u32 pack(float re, float im) {
    f32x2 v = {re, im};
    return __builtin_bit_cast(u32, __builtin_convertvector(v, f16x2));
}
__device__ __forceinline__ u32 pack(cf v) { return pack(v.x, v.y); }
// Packed fp16 complex sum / difference of two words (v_pk_add_f16): the
// exact result rounded once to fp16, the same as or closer than an FP32
// add followed by the rounding to fp16.
__device__ __forceinline__ u32 hadd2(u32 a, u32 b) {
    return __builtin_bit_cast(u32, __builtin_bit_cast(f16x2, a) + __builtin_bit_cast(f16x2, b));
}
__device__ __forceinline__ u32 hsub2(u32 a, u32 b) {
    return __builtin_bit_cast(u32, __builtin_bit_cast(f16x2, a) - __builtin_bit_cast(f16x2, b));
}
__device__ __forceinline__ cf unpack(u32 w) {
    f16x2 h = __builtin_bit_cast(f16x2, w);
    return {(float)h[0], (float)h[1]};
}

// Twiddle table entry: tables hold FP64-generated exp(-2*pi*i*m/M) as FP32;
// the adjoint direction (sign = +1) uses the conjugate.
__device__ __forceinline__ cf twiddle(const float2* __restrict__ tab, u32 idx, float sgn_im) {
    float2 t = tab[idx];
    return {t.x, sgn_im * t.y};
}

// cos(2*pi*m/16), m = 0..15 (used for the DFT-16 operand), from the
// hardware cosine (argument in revolutions, m/16 exact) rather than a table:
// a per-lane table lookup is a global load, and global loads complete in
// issue order, so it would wait behind the data loads.  The results are
// within a few FP32 ulps, far from every FP16 rounding boundary of these
// values (the nearest is 0.029 FP16 ulp away), so the FP16 operand is the
// correctly rounded DFT-16 matrix; |cos| below 2^-20 is the exact zero.
__device__ __forceinline__ float cos16(int m) {
    const float c = __builtin_amdgcn_cosf((float)(m & 15) * 0.0625f);
    return fabsf(c) < 0x1p-20f ? 0.f : c;
}
__device__ __forceinline__ float sin16(int m) { return cos16(m - 4); }

// Correctly rounded FP32 cos(2*pi*m/16) (split precision keeps the FP32
// value's low part, so the hardware cosine's last-ulp error would show).
__device__ __forceinline__ float cos16_exact(int m) {
    constexpr float c[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                             0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                             -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                             0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
    return c[m & 15];
}
__device__ __forceinline__ float sin16_exact(int m) { return cos16_exact(m - 4); }

// Register budget of the MFMA kernels launched with at most 256 threads.
// Such a workgroup may run one wave per SIMD, so without a cap each lane may
// use 512 registers (256 VGPRs + 256 AGPRs), and with that budget the
// compiler writes every MFMA result to AGPRs and copies it out with one
// v_accvgpr_read per register (4 VALU ops per MFMA).  At least 2 waves per
// SIMD caps the budget at 256 and keeps MFMA results in VGPRs.
#define TCFFT_MFMA_BUDGET __attribute__((amdgpu_waves_per_eu(2)))

struct Dft16Op {
    f16x8 re, im;
};

// Lane index within the wave.  The value passes through an empty volatile
// asm so that the compiler cannot hoist lane-derived constants (DFT-16
// operands, twiddles) out of a persistent task loop, where they would stay
// live across every task and exceed the register budget.
__device__ __forceinline__ int lane_id() {
    int l = (int)(threadIdx.x & 63u);
    asm volatile("" : "+v"(l));
    return l;
}

// Constant DFT-16 operand for this lane.  sign = -1 forward, +1 adjoint.
__device__ __forceinline__ Dft16Op make_dft16(int lane, int sign) {
    const int f = lane & 15, g = lane >> 4;
    Dft16Op op;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = 4 * g + j;
        const int m = (n * f) & 15;
        const float fr = cos16(m);
        const float fi = sign < 0 ? -sin16(m) : sin16(m);
        op.re[2 * j] = (_Float16)fr;
        op.re[2 * j + 1] = (_Float16)(-fi);
        op.im[2 * j] = (_Float16)fi;
        op.im[2 * j + 1] = (_Float16)fr;
    }
    return op;
}

// Operand of a DFT-16 whose output row f is multiplied by scale (a complex
// FP32 value, here passed as the one for this lane's row f = lane%16).
// Used in the B-form, where the constant is the A operand and its row is the
// output frequency: D = diag(scale) * F * X.
__device__ __forceinline__ Dft16Op make_dft16_scaled(int lane, int sign, cf scale) {
    const int f = lane & 15, g = lane >> 4;
    Dft16Op op;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = 4 * g + j;
        const int m = (n * f) & 15;
        const float fr0 = cos16(m);
        const float fi0 = sign < 0 ? -sin16(m) : sin16(m);
        const float fr = scale.x * fr0 - scale.y * fi0;
        const float fi = scale.x * fi0 + scale.y * fr0;
        op.re[2 * j] = (_Float16)fr;
        op.re[2 * j + 1] = (_Float16)(-fi);
        op.im[2 * j] = (_Float16)fi;
        op.im[2 * j + 1] = (_Float16)fr;
    }
    return op;
}

__device__ __forceinline__ f16x8 as_operand(u32 a, u32 b, u32 c, u32 d) {
    u32x4 v = {a, b, c, d};
    return __builtin_bit_cast(f16x8, v);
}

// A-form: data (4 words = 4 points of column lane%16) x DFT16.
__device__ __forceinline__ void dft16_a(const u32* w, const Dft16Op& op, f32x4& re, f32x4& im) {
    const f16x8 a = as_operand(w[0], w[1], w[2], w[3]);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    re = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, op.re, z, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, op.im, z, 0, 0, 0);
}

// B-form: DFT16 x data.
__device__ __forceinline__ void dft16_b(const u32* w, const Dft16Op& op, f32x4& re, f32x4& im) {
    const f16x8 b = as_operand(w[0], w[1], w[2], w[3]);
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    re = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.re, b, z, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f32_16x16x32_f16(op.im, b, z, 0, 0, 0);
}

// Swap lane bit 4 (resp. 5) with a register bit: x is the register whose bit
// is 0, y the one whose bit is 1.  After the call register bit b at lane bit a
// holds what register bit a at lane bit b held.  (v_permlane16_swap_b32 /
// v_permlane32_swap_b32, gfx950.)
__device__ __forceinline__ void swap_l4(u32& x, u32& y) {
    auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}
__device__ __forceinline__ void swap_l5(u32& x, u32& y) {
    auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}

// Register transpose used between the in-wave stages: with register index
// (t, i) -- t = 4-bit tile, i = 2-bit word -- and lane group g = lane/16,
// exchange (g, i) with t: afterwards (g, i) holds what t held and t holds
// (old i in bits 0-1, old g in bits 2-3).  A rename plus two permlane rounds.
__device__ __forceinline__ void transpose_gi_tile(u32 (&src)[16][4], u32 (&dst)[16][4]) {
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(t & 12) | i][t & 3] = src[t][i];
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if (!(t & 4))
#pragma unroll
            for (int i = 0; i < 4; ++i) swap_l4(dst[t][i], dst[t | 4][i]);
#pragma unroll
    for (int t = 0; t < 16; ++t)
        if (!(t & 8))
#pragma unroll
            for (int i = 0; i < 4; ++i) swap_l5(dst[t][i], dst[t | 8][i]);
}

// The same transpose on FP32 data (register renames and lane swaps only).
__device__ __forceinline__ void transpose_gi_tile(float (&src)[16][4], float (&dst)[16][4]) {
    u32 a[16][4], b[16][4];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) a[t][i] = __float_as_uint(src[t][i]);
    transpose_gi_tile(a, b);
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[t][i] = __uint_as_float(b[t][i]);
}

// Streaming (non-temporal) variants for data touched once per pass.
__device__ __forceinline__ void load4_nt(u32* w, const u32* __restrict__ p) {
    u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    w[0] = v[0];
    w[1] = v[1];
    w[2] = v[2];
    w[3] = v[3];
}
__device__ __forceinline__ void store4_nt(u32* __restrict__ p, const u32* w) {
    u32x4 v = {w[0], w[1], w[2], w[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
}

__device__ __forceinline__ void load4(u32* w, const u32* __restrict__ p) {
    u32x4 v = *reinterpret_cast<const u32x4*>(p);
    w[0] = v[0];
    w[1] = v[1];
    w[2] = v[2];
    w[3] = v[3];
}
__device__ __forceinline__ void store4(u32* __restrict__ p, const u32* w) {
    u32x4 v = {w[0], w[1], w[2], w[3]};
    *reinterpret_cast<u32x4*>(p) = v;
}

// Access at a wave-uniform base plus a 32-bit per-lane byte offset: the
// compiler then uses the SGPR-base form of global loads and stores (base in
// SGPRs, offset in one VGPR) instead of a 64-bit address per lane, which
// for a strided pass costs three VALU ops and two VGPRs per access.
template <class T>
__device__ __forceinline__ T ld_at(const void* base, u32 off_bytes) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off_bytes);
}
template <class T>
__device__ __forceinline__ T ld_at_nt(const void* base, u32 off_bytes) {
    return __builtin_nontemporal_load(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off_bytes));
}
template <class T>
__device__ __forceinline__ void st_at(void* base, u32 off_bytes, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off_bytes) = v;
}
template <class T>
__device__ __forceinline__ void st_at_nt(void* base, u32 off_bytes, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off_bytes));
}

// Small radix-T DFT in FP32 registers (T = 2, 4, 8), in place on v[0..T).
template <int T>
__device__ __forceinline__ void small_dft(cf* v, int sign);

template <>
__device__ __forceinline__ void small_dft<1>(cf*, int) {}

template <>
__device__ __forceinline__ void small_dft<2>(cf* v, int) {
    cf a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

// multiply by sign*i  (exp(sign * i*pi/2))
__device__ __forceinline__ cf mul_si(cf a, int sign) {
    return sign < 0 ? cf{a.y, -a.x} : cf{-a.y, a.x};
}

template <>
__device__ __forceinline__ void small_dft<4>(cf* v, int sign) {
    cf s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cf s13 = cadd(v[1], v[3]), d13 = mul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <>
__device__ __forceinline__ void small_dft<8>(cf* v, int sign) {
    // even/odd split: X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
    cf e[4] = {v[0], v[2], v[4], v[6]};
    cf o[4] = {v[1], v[3], v[5], v[7]};
    small_dft<4>(e, sign);
    small_dft<4>(o, sign);
    const float h = 0.70710678118654752f;
    // W8^1 = (1, sign)*h, W8^2 = sign*i, W8^3 = (-1, sign)*h
    cf w1 = cmul(o[1], cf{h, sign * h});
    cf w2 = mul_si(o[2], sign);
    cf w3 = cmul(o[3], cf{-h, sign * h});
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
}

}  // namespace tcfft
