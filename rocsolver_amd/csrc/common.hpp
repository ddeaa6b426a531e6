// common.hpp -- shared device/host helpers for the MI355X (gfx950) batched LU engine.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <type_traits>
#include <utility>

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

    // Every batch entry is device global memory.  The cast through the
    // global address space says so to the compiler, so accesses through the
    // returned pointer are global_* instructions (SGPR base + VGPR offset,
    // counted by vmcnt only) instead of flat ones, which also hold up every
    // wait on LDS.
    __device__ __forceinline__ T* operator[](int64_t b) const
    {
        T* p = ptrs ? ptrs[b] : base + b * stride;
        return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
    }
};

// A column-major matrix (leading dimension ld) seen through a buffer
// resource.  A load takes a per-lane byte offset (one VGPR) and a uniform
// byte offset (SGPR), so a thread walking a column block needs no address
// registers beyond its own offset; a per-lane offset of OUT_OF_RANGE returns
// 0 without touching memory.  The matrix must span fewer than 2^31 bytes
// (fits_buffer on the host).
template <typename T>
struct MatrixBuffer
{
    static constexpr uint32_t OUT_OF_RANGE = 0x80000000u;
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t ld_bytes;

    __device__ __forceinline__ MatrixBuffer(const T* base, int ld, int ncols)
        : MatrixBuffer(base, ld, uint64_t(ld) * uint64_t(ncols) * sizeof(T))
    {
    }
    // a buffer of `bytes` bytes from base (several matrices of one wavefront)
    __device__ __forceinline__ MatrixBuffer(const T* base, int ld, uint64_t bytes)
    {
        // word 3 of a raw gfx9 buffer resource: 32-bit data format, no swizzle
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0,
                                                 int(bytes < OUT_OF_RANGE ? bytes : OUT_OF_RANGE - 1),
                                                 0x00020000);
        ld_bytes = uint32_t(ld) * uint32_t(sizeof(T));
    }
    __device__ __forceinline__ uint32_t row_offset(int i) const
    {
        return uint32_t(i) * uint32_t(sizeof(T));
    }
    __device__ __forceinline__ uint32_t col_offset(int j) const
    {
        return uint32_t(j) * ld_bytes;
    }
    __device__ __forceinline__ T load(uint32_t voff, uint32_t soff) const
    {
        if constexpr(sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, soff, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, soff, 0));
    }
    // a store at an out-of-range offset is dropped by the hardware
    __device__ __forceinline__ void store(T v, uint32_t voff, uint32_t soff) const
    {
        if constexpr(sizeof(T) == 8)
        {
            using bits_t = decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc, 0, 0, 0));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bits_t, v), rsrc, voff, soff, 0);
        }
        else
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rsrc, voff, soff, 0);
    }
    // the byte offset (row i, column j) for an active lane, else out of range
    __device__ __forceinline__ uint32_t offset(bool active, int i, int j) const
    {
        return active ? row_offset(i) + col_offset(j) : OUT_OF_RANGE;
    }
};

// A value equal in every lane, moved to scalar registers (a uniform value
// the compiler cannot prove uniform would otherwise occupy vector registers
// for as long as it lives).
__device__ __forceinline__ double uniform_value(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// threadIdx.x, opaque to the optimiser: addresses derived from it are then
// computed where they are used instead of once at kernel entry and kept live
// (in registers, or spilled) through every phase of a long kernel.
__device__ __forceinline__ int fresh_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Stops the optimiser from precomputing a sequence of running offsets (it
// would keep every element of the sequence live in its own register).
__device__ __forceinline__ void opaque(uint32_t& v)
{
    asm volatile("" : "+v"(v));
}

template <typename T>
inline bool fits_buffer(int ld, int ncols)
{
    return uint64_t(ld) * uint64_t(ncols) * sizeof(T) < (uint64_t(1) << 31);
}

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

// Value of v in lane l (l uniform across the wavefront) through v_readlane:
// a few cycles, no LDS.
__device__ __forceinline__ float readlane_t(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double readlane_t(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Orders LDS accesses between the lanes of one wavefront.  A wavefront's
// LDS instructions execute in issue order, so this only has to stop the
// compiler from moving memory operations across it (no s_barrier: teams of
// lanes that never span two waves synchronise with this alone, even when the
// other waves of the workgroup run different code).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double fma_t(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fma_t(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}

// correctly rounded square roots (the IEEE operation, as the host's)
__device__ __forceinline__ double sqrt_t(double a)
{
    return __builtin_sqrt(a);
}
__device__ __forceinline__ float sqrt_t(float a)
{
    return __builtin_sqrtf(a);
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

// Largest signed / unsigned 32-bit integer over the group (every lane gets
// it); the compiler folds each DPP move into the max (v_max_*_dpp).
template <int TS, typename I>
__device__ __forceinline__ I group_max_int(I v)
{
    auto mx = [](I a, I b) { return a > b ? a : b; };
    v = mx(v, I(dpp_mov<0xB1>(int(v))));
    v = mx(v, I(dpp_mov<0x4E>(int(v))));
    v = mx(v, I(dpp_mov<0x141>(int(v))));
    v = mx(v, I(dpp_mov<0x140>(int(v))));
    if(TS >= 32)
    {
        lane_pair p = swap16(int(v));
        v = mx(I(p.a), I(p.b));
    }
    if(TS >= 64)
    {
        lane_pair p = swap32(int(v));
        v = mx(I(p.a), I(p.b));
    }
    return v;
}

// Arg-max of (key, index) over a group of TS lanes, ties to the smaller
// index; every lane of the group ends with the winner.  Keys are pivot_key
// values: non-negative (NaN already mapped to +inf) or -1 for a lane that is
// not a candidate.  The IEEE bit patterns of non-negative values order like
// the values, so the maximum is found on integers -- for double the high
// words first, then the low words among the lanes that hold the largest high
// word -- with DPP folded into each integer max, where a floating-point max
// would need a separate DPP move and NaN canonicalisation per step.
// Non-candidates enter with key bits 0 and lose every tie to a candidate
// (index tagged with bit 30); with no candidate at all the result is the
// smallest index and key -1, as a maximum over the -1 keys would give (an
// index of INT_MAX stands for "no row" and stays INT_MAX).
// Indices must be below 2^30 (or INT_MAX for "none").
template <int TS>
__device__ __forceinline__ void group_argmax(double& key, int& idx)
{
    const bool cand = key >= 0.0;
    const uint64_t u = cand ? uint64_t(__double_as_longlong(key)) : 0u;
    const int hi = int(u >> 32); // < 2^31: the sign bit is clear
    const unsigned lo = unsigned(u);
    const int mh = group_max_int<TS>(hi);
    const unsigned ml = group_max_int<TS>(hi == mh ? lo : 0u);
    constexpr int TAG = 0x40000000;
    const int r = group_min<TS>((hi == mh && lo == ml) ? (cand ? idx : (idx | TAG)) : 0x7fffffff);
    const bool none = r == 0x7fffffff || (r & TAG) != 0;
    idx = (r == 0x7fffffff) ? r : (r & ~TAG);
    key = none ? -1.0 : __longlong_as_double((long long)((uint64_t(unsigned(mh)) << 32) | ml));
}

template <int TS>
__device__ __forceinline__ void group_argmax(float& key, int& idx)
{
    const bool cand = key >= 0.f;
    const int u = cand ? __float_as_int(key) : 0; // < 2^31: the sign bit is clear
    const int m = group_max_int<TS>(u);
    constexpr int TAG = 0x40000000;
    const int r = group_min<TS>(u == m ? (cand ? idx : (idx | TAG)) : 0x7fffffff);
    const bool none = r == 0x7fffffff || (r & TAG) != 0;
    idx = (r == 0x7fffffff) ? r : (r & ~TAG);
    key = none ? -1.f : __int_as_float(m);
}

template <typename T>
__device__ __forceinline__ void wave_argmax(T& key, int& idx)
{
    group_argmax<64>(key, idx);
}

// Column maxima over the 64 lanes of a wavefront of K values per lane (K a
// power of two, 2 <= K <= 64), by halving exchanges: at each step a lane
// keeps half of its columns and hands the other half to the partner lane
// (lane ^ 32, then ^ 16, ...), so the steps take K/2 + K/4 + ... + 1
// exchanges instead of K full reductions; the 64/K lanes that then share a
// column reduce it.  Every lane returns the maximum (fmax: NaN ignored) of
// column lane / (64 / K).
template <int K>
__device__ __forceinline__ double wave_colmax(const double (&v)[K])
{
    static_assert(K >= 2 && K <= 64 && (K & (K - 1)) == 0, "K: power of two in [2, 64]");
    const int lane = threadIdx.x & 63;
    double w[K];
#pragma unroll
    for(int j = 0; j < K; ++j)
        w[j] = v[j];
#pragma unroll
    for(int half = K / 2, bit = 32; half >= 1; half /= 2, bit /= 2)
    {
        const bool up = (lane & bit) != 0; // keeps the upper half of its columns
#pragma unroll
        for(int j = 0; j < half; ++j)
        {
            const double keep = up ? w[half + j] : w[j];
            const double send = up ? w[j] : w[half + j];
            w[j] = fmax(keep, __shfl_xor(send, bit));
        }
    }
#pragma unroll
    for(int b = 32 / K; b >= 1; b /= 2)
        w[0] = fmax(w[0], __shfl_xor(w[0], b));
    return w[0];
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
