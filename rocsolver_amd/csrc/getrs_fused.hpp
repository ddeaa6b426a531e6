// getrs_fused.hpp -- one-workgroup-per-system triangular solves with the LU
// of getrf, for SMALL_MAX < n <= FUSED_MAX (reference order of operations:
// library/src/lapack/roclapack_getrs.hpp:117-150 -- laswp, then L, then U for
// op = N; U^T, L^T, then the reverse interchanges for op = T).
//
// Thread i owns row i of NR right-hand sides in registers; the triangles are
// walked by the row solvers of trsv_rows.hpp (32-row diagonal blocks, loads
// one block ahead, one barrier per block).  The factors are read through a
// buffer resource: lda * n * sizeof(TL) must be below 2^31 (fits_buffer).
//
// Types: TL is the LU precision and the arithmetic precision; TB is the
// precision of the right-hand side read from BB and of the solution written
// to XX (XX may alias BB).  ACCUM adds the result to XX instead of storing it
// (the correction step of iterative refinement).  Systems with skip[b] != 0
// are left untouched.
#pragma once

#include "internal.hpp"
#include "trsv_rows.hpp"

#include <type_traits>

namespace rsa
{

// Solves one system with the whole workgroup (blockDim.x == NMAX, n <= NMAX).
template <typename TL, typename TB, int NMAX, int NR, bool ACCUM, bool TRANS>
__device__ __forceinline__ void getrs_fused_body(int n, int nrhs, const TL* __restrict__ Ap,
                                                 int lda, const int* __restrict__ ipiv,
                                                 const TB* Bm, int ldb, TB* X, int ldx)
{
    __shared__ TL sx[2][GBS][NR];
    __shared__ TL sL[GBS][GBS + 1];
    __shared__ int sipiv[NMAX];
    __shared__ int ssrc[NMAX];

    const int tid = threadIdx.x;
    const MatrixBuffer<TL> A(Ap, lda, n);
    const MatrixBuffer<TB> Bbuf(Bm, ldb, nrhs);
    const int pos = pivot_position(n, ipiv, TRANS, sipiv, ssrc);
    const int src = TRANS ? tid : ssrc[tid];
    const int dst = TRANS ? pos : tid;

    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        // all NR loads issued back to back (out-of-range offsets read 0)
        TL x[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
        {
            const uint32_t off = (tid < n && c0 + q < nrhs)
                                     ? Bbuf.row_offset(src) + Bbuf.col_offset(c0 + q)
                                     : MatrixBuffer<TB>::OUT_OF_RANGE;
            x[q] = TL(Bbuf.load(off, 0));
        }
        solve_lower_rows<TL, NR, TRANS>(n, A, x, sx, sL);
        solve_upper_rows<TL, NR, TRANS>(n, A, x, sx, sL);
        // every thread has read its right-hand side before any row is stored
        // (X may alias B and op = T stores at the permuted row)
        __syncthreads();
        if(tid < n)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                if(c0 + q < nrhs)
                {
                    TB& o = X[dst + int64_t(c0 + q) * ldx];
                    o = ACCUM ? TB(o + TB(x[q])) : TB(x[q]);
                }
        }
    }
}

template <typename TL, typename TB, int NMAX, int NR, bool ACCUM, bool TRANS>
__global__ __launch_bounds__(NMAX) void getrs_fused_kernel(int n, int nrhs, Batch<TL> AA, int lda,
                                                           const int* __restrict__ ipivA,
                                                           int64_t strideP, Batch<TB> BB, int ldb,
                                                           Batch<TB> XX, int ldx,
                                                           const int* __restrict__ skip)
{
    const int64_t b = blockIdx.x;
    if(skip && skip[b] != 0)
        return;
    getrs_fused_body<TL, TB, NMAX, NR, ACCUM, TRANS>(n, nrhs, AA[b], lda, ipivA + b * strideP,
                                                     BB[b], ldb, XX[b], ldx);
}

template <typename TL, typename TB, bool ACCUM>
rocblas_status getrs_fused_launch(hipStream_t stream, bool trans, int n, int nrhs, Batch<TL> A,
                                  int lda, const int* ipiv, int64_t strideP, Batch<TB> B, int ldb,
                                  Batch<TB> X, int ldx, int bc, const int* skip)
{
    if(bc <= 0 || n <= 0 || nrhs <= 0)
        return rocblas_status_success;
    auto go = [&](auto nmax, auto nr, auto tr) {
        constexpr int NMAX = decltype(nmax)::value;
        constexpr int NR = decltype(nr)::value;
        constexpr bool TR = decltype(tr)::value;
        hipLaunchKernelGGL((getrs_fused_kernel<TL, TB, NMAX, NR, ACCUM, TR>), dim3(bc), dim3(NMAX),
                           0, stream, n, nrhs, A, lda, ipiv, strideP, B, ldb, X, ldx, skip);
    };
    auto by_trans = [&](auto nmax, auto nr) {
        if(trans)
            go(nmax, nr, std::true_type{});
        else
            go(nmax, nr, std::false_type{});
    };
    auto by_nrhs = [&](auto nmax) {
        if(nrhs <= 1)
            by_trans(nmax, std::integral_constant<int, 1>{});
        else if(nrhs <= 4)
            by_trans(nmax, std::integral_constant<int, 4>{});
        else
            by_trans(nmax, std::integral_constant<int, 8>{});
    };
    if(n <= 128)
        by_nrhs(std::integral_constant<int, 128>{});
    else if(n <= 256)
        by_nrhs(std::integral_constant<int, 256>{});
    else
        by_nrhs(std::integral_constant<int, 512>{});
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

} // namespace rsa
