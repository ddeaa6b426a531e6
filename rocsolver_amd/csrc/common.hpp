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

// ----------------------------------------------------------------------------
// Cross-lane reductions over groups of TS = 16, 32 or 64 lanes without LDS:
// DPP quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and row_mirror reduce
// within each 16-lane row; v_permlane16_swap pairs rows (0,1) and (2,3) and
// v_permlane32_swap the two halves of the wave.  With both operands equal to
// x, the two results of a swap hold x and the partner row's x in every lane,
// so max/min of the pair is the pairwise reduction.
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

struct lane_pair
{
    int a, b;
};
__device__ __forceinline__ lane_pair swap16(int v)
{
    auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return {int(p[0]), int(p[1])};
}
__device__ __forceinline__ lane_pair swap32(int v)
{
    auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return {int(p[0]), int(p[1])};
}

// maximum of a float/double over the group (every lane gets it)
template <int TS>
__device__ __forceinline__ float group_max(float v)
{
    v = fmaxf(v, __int_as_float(dpp_mov<0xB1>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_mov<0x4E>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_mov<0x141>(__float_as_int(v))));
    v = fmaxf(v, __int_as_float(dpp_mov<0x140>(__float_as_int(v))));
    if(TS >= 32)
    {
        lane_pair p = swap16(__float_as_int(v));
        v = fmaxf(__int_as_float(p.a), __int_as_float(p.b));
    }
    if(TS >= 64)
    {
        lane_pair p = swap32(__float_as_int(v));
        v = fmaxf(__int_as_float(p.a), __int_as_float(p.b));
    }
    return v;
}

__device__ __forceinline__ double dbl_from(int lo, int hi)
{
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v)
{
    return dbl_from(dpp_mov<CTRL>(__double2loint(v)), dpp_mov<CTRL>(__double2hiint(v)));
}

template <int TS>
__device__ __forceinline__ double group_max(double v)
{
    v = fmax(v, dpp_mov_d<0xB1>(v));
    v = fmax(v, dpp_mov_d<0x4E>(v));
    v = fmax(v, dpp_mov_d<0x141>(v));
    v = fmax(v, dpp_mov_d<0x140>(v));
    if(TS >= 32)
    {
        lane_pair lo = swap16(__double2loint(v)), hi = swap16(__double2hiint(v));
        v = fmax(dbl_from(lo.a, hi.a), dbl_from(lo.b, hi.b));
    }
    if(TS >= 64)
    {
        lane_pair lo = swap32(__double2loint(v)), hi = swap32(__double2hiint(v));
        v = fmax(dbl_from(lo.a, hi.a), dbl_from(lo.b, hi.b));
    }
    return v;
}

template <int TS>
__device__ __forceinline__ int group_min(int v)
{
    v = min(v, dpp_mov<0xB1>(v));
    v = min(v, dpp_mov<0x4E>(v));
    v = min(v, dpp_mov<0x141>(v));
    v = min(v, dpp_mov<0x140>(v));
    if(TS >= 32)
    {
        lane_pair p = swap16(v);
        v = min(p.a, p.b);
    }
    if(TS >= 64)
    {
        lane_pair p = swap32(v);
        v = min(p.a, p.b);
    }
    return v;
}

// Arg-max of (key, index) over a group of TS lanes, ties to the smaller
// index; every lane of the group ends with the winner.  Keys must not be
// NaN (pivot_key maps NaN to +-inf / -1).
template <int TS, typename T>
__device__ __forceinline__ void group_argmax(T& key, int& idx)
{
    const T m = group_max<TS>(key);
    idx = group_min<TS>(key == m ? idx : 0x7fffffff);
    key = m;
}

template <typename T>
__device__ __forceinline__ void wave_argmax(T& key, int& idx)
{
    group_argmax<64>(key, idx);
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
