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

} // namespace rsa
