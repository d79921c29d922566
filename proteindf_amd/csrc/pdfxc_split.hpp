// bf16 slicing and the bf16 MFMA shared by the split-precision kernels
// (pdfxc_split.hip) and the split XC gradient (pdfxc_grad.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace pdfxc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16 (checked on gfx950 by pdfxc_selftest_mfma_split):
// lane l holds A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16],
// e = 0..7; register r of the result holds D[4 (l / 16) + r][l % 16].
__device__ inline f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline unsigned short bf16_bits(__bf16 h) { return __builtin_bit_cast(unsigned short, h); }

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_f16: the lane layout of mfma_bf16 (checked by the same
// self-test)
__device__ inline f32x4 mfma_f16(f16x8 a, f16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x16_f16: lane l holds A[l % 16][4 (l / 16) + e] and
// B[4 (l / 16) + e][l % 16], e = 0..3; the result as for the 16x16x32 forms
__device__ inline f32x4 mfma_f16_k16(f16x4 a, f16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Keeps a float as computed: without it the compiler folds a double ->
// float -> bf16 / fp16 pair of conversions into one double -> bf16 / fp16
// conversion, which gfx950 has no instruction for (7 extra VALU per bf16
// value, about 35 per fp16 value); the hardware float -> bf16 / fp16
// conversions then round the float
__device__ inline float keep_f32(float f) {
    __asm__ volatile("" : "+v"(f));
    return f;
}

// FP16 hi + lo of t, |t| <= 2^15: hi = f16(float(t)), lo = f16(float(t - hi))
// (t - hi is exact in FP64); the pair keeps 22 bits of t down to |t| ~ 2^-3,
// and an absolute 2^-25 below that (FP16 subnormals)
__device__ inline void split_f16(double t, unsigned short* out) {
    const _Float16 h = (_Float16)keep_f32((float)t);
    out[0] = __builtin_bit_cast(unsigned short, h);
    out[1] = __builtin_bit_cast(unsigned short, (_Float16)keep_f32((float)(t - (double)(float)h)));
}

// slices of x: s[0] = bf16(f), f = float(x), s[1] = bf16 of what s[0] leaves, and the last
// slice from what the others leave plus e = float(x - f), the FP64 part that
// f = float(x) drops.  f - s[0] and r - s[1] are exact in FP32 (each slice
// keeps the top 8 bits of a value with at most 24), so only the last sum
// rounds, at 2^-24 of a value below 2^-16 |x|; one FP64 subtraction per x.
template <int S>
__device__ inline void split_bf16(double x, unsigned short* out) {
    const float f = keep_f32((float)x);
    const float e = (float)(x - (double)f);
    const __bf16 h0 = (__bf16)f;
    out[0] = bf16_bits(h0);
    const float r1 = f - (float)h0;  // exact
    if constexpr (S == 2) {
        out[1] = bf16_bits((__bf16)(r1 + e));
    } else {
        const __bf16 h1 = (__bf16)r1;
        out[1] = bf16_bits(h1);
        out[2] = bf16_bits((__bf16)((r1 - (float)h1) + e));  // r1 - h1 exact
    }
}

// slice products kept: (a, b) with a + b < S
template <int S>
struct SlicePairs;
template <>
struct SlicePairs<2> {
    static constexpr int n = 3;
    static constexpr int a[3] = {0, 0, 1};
    static constexpr int b[3] = {0, 1, 0};
};
template <>
struct SlicePairs<3> {
    static constexpr int n = 6;
    static constexpr int a[6] = {0, 0, 1, 0, 2, 1};
    static constexpr int b[6] = {0, 1, 0, 2, 0, 1};
};

}  // namespace pdfxc
