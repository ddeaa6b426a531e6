// common.hpp -- shared device/host helpers for the MI355X (gfx950) batched LU engine.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>

namespace rsa
{

// Batch addressing.  Strided form: base + b*stride; batched form: ptrs[b]
// (a device array of device pointers).  Mirrors the reference's
// load_ptr_batch (common/include/rocblas_utility.hpp:97-117).
template <typename T>
struct Batch
{
    T* base;
    T* const* ptrs;
    int64_t stride;

    __device__ __forceinline__ T* operator[](int64_t b) const
    {
        return ptrs ? ptrs[b] : base + b * stride;
    }
};

template <typename T>
inline Batch<T> strided(T* base, int64_t stride)
{
    return Batch<T>{base, nullptr, stride};
}

template <typename T>
inline Batch<T> arrayed(T* const* ptrs)
{
    return Batch<T>{nullptr, ptrs, 0};
}

__device__ __forceinline__ double fma_t(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fma_t(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

__device__ __forceinline__ double abs_t(double a)
{
    return __builtin_fabs(a);
}
__device__ __forceinline__ float abs_t(float a)
{
    return __builtin_fabsf(a);
}

template <typename T>
__device__ __forceinline__ bool isnan_t(T a)
{
    return a != a;
}

// Pivot-candidate key.  The reference scan (first index of max |a| with a
// strict '<' comparison, roclapack_getf2_specialized_kernels.hpp:85-93)
// never lets a NaN displace the candidate and keeps the first row (k) when
// that row itself is NaN.  As a max-reduction with ties to the lowest row:
// NaN on row k -> +inf, NaN elsewhere -> -1 (never chosen), else |a|.
template <typename T>
__device__ __forceinline__ T pivot_key(T a, bool is_first_row)
{
    if(isnan_t(a))
        return is_first_row ? T(__builtin_inf()) : T(-1);
    return abs_t(a);
}

// Arg-max of (key, index) over the whole wavefront, ties to the smaller
// index; every lane ends with the winner.
template <typename T>
__device__ __forceinline__ void wave_argmax(T& key, int& idx)
{
#pragma unroll
    for(int off = 1; off < 64; off <<= 1)
    {
        T ok = __shfl_xor(key, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        if(ok > key || (ok == key && oi < idx))
        {
            key = ok;
            idx = oi;
        }
    }
}

inline int ceil_div(int64_t a, int64_t b)
{
    return int((a + b - 1) / b);
}

} // namespace rsa

#define RSA_LAUNCH_CHECK()                                 \
    do                                                     \
    {                                                      \
        if(hipGetLastError() != hipSuccess)                \
            return rocblas_status_internal_error;          \
    } while(0)
