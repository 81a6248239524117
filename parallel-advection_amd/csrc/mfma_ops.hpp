// Operand types and conversions of the MFMA Conv1d kernels (gfx950).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

namespace bkma {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }  // round to nearest even

// Two FP32 values -> two bf16 (round to nearest even), a in the low half:
// one v_cvt_pk_bf16_f32.
__device__ __forceinline__ unsigned pack_bf16x2(float a, float b) {
    typedef float f32x2_v __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2_v __attribute__((ext_vector_type(2)));
    const f32x2_v v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_v));
}

// Largest finite OCP FP8 e4m3 value.
constexpr float kFp8E4m3Max = 448.0f;

// |v| beyond 448 (infinities included) becomes +-448; NaN stays NaN.  The
// hardware conversion below has no saturation of its own: it turns values
// that round beyond 448 into the NaN code.
__device__ __forceinline__ float fp8_saturate(float v) {
    return fabsf(v) > kFp8E4m3Max ? copysignf(kFp8E4m3Max, v) : v;
}

// Two FP32 values -> two OCP e4m3 bytes in the low half (round to nearest
// even, saturating to +-448, NaN kept).
__device__ __forceinline__ int pack_fp8x2(float a, float b) {
    return __builtin_amdgcn_cvt_pk_fp8_f32(fp8_saturate(a), fp8_saturate(b), 0, false);
}

}  // namespace bkma
