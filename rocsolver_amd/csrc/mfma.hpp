// mfma.hpp -- 16x16x4 matrix-core helpers (gfx950).
//
// Operand layout of v_mfma_{f32,f64}_16x16x4 (lane l of the wave64):
//   A (16x4):  lane l supplies A[l%16][l/16]
//   B (4x16):  lane l supplies B[l/16][l%16]
//   D (16x16): result register r of lane l holds column l%16 and row
//              4*(l/16)+r for the f32 instruction, but row l/16+4*r for the
//              f64 instruction (measured with a one-instruction probe on
//              gfx950); mfma16<T>::row(lane, r) returns it.
#pragma once

#include <hip/hip_runtime.h>

namespace rsa
{

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct mfma16;

template <>
struct mfma16<double>
{
    typedef d4_t acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r)
    {
        return lane / 16 + 4 * r;
    }
};

template <>
struct mfma16<float>
{
    typedef f4_t acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r)
    {
        return 4 * (lane / 16) + r;
    }
};

// ---- 16x16x16 with half-width operands and FP32 accumulation
//   A (16x16): lane l supplies A[l%16][4*(l/16) + e], e = 0..3
//   B (16x16): lane l supplies B[4*(l/16) + e][l%16]
//   D (16x16): as v_mfma_f32_16x16x4f32 (row 4*(l/16)+r, column l%16)
// Operand kinds: MFMA_F16 (IEEE binary16, v_mfma_f32_16x16x16_f16) and
// MFMA_BF16 (bfloat16, v_mfma_f32_16x16x16_bf16); MFMA_NATIVE means the
// 16x16x4 instruction of the factor precision itself.
enum MfmaOperand : int
{
    MFMA_NATIVE = 0,
    MFMA_F16 = 1,
    MFMA_BF16 = 2
};

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef short s4_t __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (NaN stays a quiet NaN)
__device__ __forceinline__ short bf16_bits(float x)
{
    unsigned u = __float_as_uint(x);
    if((u & 0x7fffffffu) > 0x7f800000u)
        return short((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

template <int OPK>
__device__ __forceinline__ f4_t mfma16x16x16(const float (&a)[4], const float (&b)[4], f4_t c)
{
    if constexpr(OPK == MFMA_F16)
    {
        const h4_t ah = {_Float16(a[0]), _Float16(a[1]), _Float16(a[2]), _Float16(a[3])};
        const h4_t bh = {_Float16(b[0]), _Float16(b[1]), _Float16(b[2]), _Float16(b[3])};
        return __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, c, 0, 0, 0);
    }
    else
    {
        const s4_t ab = {bf16_bits(a[0]), bf16_bits(a[1]), bf16_bits(a[2]), bf16_bits(a[3])};
        const s4_t bb = {bf16_bits(b[0]), bf16_bits(b[1]), bf16_bits(b[2]), bf16_bits(b[3])};
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ab, bb, c, 0, 0, 0);
    }
}

} // namespace rsa
