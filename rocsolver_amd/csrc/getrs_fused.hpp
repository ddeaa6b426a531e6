// getrs_fused.hpp -- one-workgroup-per-system triangular solves with the LU
// of getrf, for SMALL_MAX < n <= FUSED_MAX (reference order of operations:
// library/src/lapack/roclapack_getrs.hpp:117-150 -- laswp, then L, then U for
// op = N; U^T, L^T, then the reverse interchanges for op = T).
//
// Thread i owns row i of NR right-hand sides in registers.  Each triangle is
// walked in 32-row diagonal blocks: the 32 threads of the block (one half
// wave) solve it with shuffles (lane k's value reaches the others without
// LDS or a barrier), publish it to LDS, and every thread below (above)
// subtracts its L (U) row segment times the block -- one barrier per block.
// Reads of the LU are coalesced across threads for op = N.
//
// Types: TL is the LU precision and the arithmetic precision; TB is the
// precision of the right-hand side read from BB and of the solution written
// to XX (XX may alias BB).  ACCUM adds the result to XX instead of storing it
// (the correction step of iterative refinement).  Systems with skip[b] != 0
// are left untouched.
#pragma once

#include "internal.hpp"

#include <type_traits>

namespace rsa
{

constexpr int GBS = 32; // diagonal block

template <typename TL, typename TB, int NMAX, int NR, bool ACCUM>
__global__ __launch_bounds__(NMAX) void getrs_fused_kernel(bool trans, int n, int nrhs,
                                                           Batch<TL> AA, int lda,
                                                           const int* __restrict__ ipivA,
                                                           int64_t strideP, Batch<TB> BB, int ldb,
                                                           Batch<TB> XX, int ldx,
                                                           const int* __restrict__ skip)
{
    __shared__ TL sx[2][GBS][NR];
    __shared__ int sipiv[NMAX];
    __shared__ int ssrc[NMAX];

    const int64_t b = blockIdx.x;
    if(skip && skip[b] != 0)
        return;
    const int tid = threadIdx.x;
    const int lane32 = tid & (GBS - 1);
    const TL* A = AA[b];
    const TB* Bm = BB[b];
    TB* X = XX[b];
    const int* ipiv = ipivA + b * strideP;
    auto L = [&](int i, int k) { return A[i + int64_t(k) * lda]; };

    // final position of original row tid after the interchanges (op = N order,
    // or reverse order for op = T); a pivot outside [0, n) is ignored
    int p_in = tid < n ? ipiv[tid] - 1 : tid;
    if(p_in < 0 || p_in >= n)
        p_in = tid;
    sipiv[tid] = p_in;
    __syncthreads();
    int pos = tid;
    if(!trans)
    {
        for(int k = 0; k < n; ++k)
        {
            const int p = sipiv[k];
            pos = (pos == k) ? p : (pos == p ? k : pos);
        }
        ssrc[pos] = tid; // row landing on `pos` comes from original row tid
    }
    else
    {
        for(int k = n - 1; k >= 0; --k)
        {
            const int p = sipiv[k];
            pos = (pos == k) ? p : (pos == p ? k : pos);
        }
    }
    __syncthreads();
    const int src = trans ? tid : ssrc[tid];
    const int nblk = (n + GBS - 1) / GBS;

    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        TL x[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
            x[q] = (tid < n && c0 + q < nrhs) ? TL(Bm[src + int64_t(c0 + q) * ldb]) : TL(0);
        int par = 0;

        // ---- first triangle: L (op = N, unit) or U^T (op = T, non-unit), forward
        for(int blk = 0; blk < nblk; ++blk, par ^= 1)
        {
            const int r0 = blk * GBS;
            if((tid >> 5) == blk) // the half wave owning rows [r0, r0+32)
            {
#pragma unroll
                for(int k = 0; k < GBS; ++k)
                {
                    if(r0 + k < n)
                    {
                        if(trans && lane32 == k)
                        {
#pragma unroll
                            for(int q = 0; q < NR; ++q)
                                x[q] = x[q] / L(r0 + k, r0 + k);
                        }
                        const TL lik = (tid < n && lane32 > k)
                                           ? (trans ? L(r0 + k, tid) : L(tid, r0 + k))
                                           : TL(0);
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                        {
                            const TL xk = __shfl(x[q], k, GBS);
                            if(lane32 > k)
                                x[q] = fma_t(-lik, xk, x[q]);
                        }
                    }
                }
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    sx[par][lane32][q] = x[q];
            }
            __syncthreads();
            if(tid >= r0 + GBS && tid < n)
            {
                const int kend = (n - r0) < GBS ? (n - r0) : GBS;
                for(int k = 0; k < kend; ++k)
                {
                    const TL lik = trans ? L(r0 + k, tid) : L(tid, r0 + k);
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-lik, sx[par][k][q], x[q]);
                }
            }
        }

        // ---- second triangle: U (op = N, non-unit) or L^T (op = T, unit), backward
        for(int blk = nblk - 1; blk >= 0; --blk, par ^= 1)
        {
            const int r0 = blk * GBS;
            if((tid >> 5) == blk)
            {
#pragma unroll
                for(int k = GBS - 1; k >= 0; --k)
                {
                    if(r0 + k < n)
                    {
                        if(!trans && lane32 == k)
                        {
#pragma unroll
                            for(int q = 0; q < NR; ++q)
                                x[q] = x[q] / L(r0 + k, r0 + k);
                        }
                        const TL uik = (lane32 < k) ? (trans ? L(r0 + k, tid) : L(tid, r0 + k))
                                                    : TL(0);
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                        {
                            const TL xk = __shfl(x[q], k, GBS);
                            if(lane32 < k)
                                x[q] = fma_t(-uik, xk, x[q]);
                        }
                    }
                }
#pragma unroll
                for(int q = 0; q < NR; ++q)
                    sx[par][lane32][q] = x[q];
            }
            __syncthreads();
            if(tid < r0)
            {
                const int kend = (n - r0) < GBS ? (n - r0) : GBS;
                for(int k = 0; k < kend; ++k)
                {
                    const TL uik = trans ? L(r0 + k, tid) : L(tid, r0 + k);
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-uik, sx[par][k][q], x[q]);
                }
            }
        }

        // ---- store (op = T: apply the interchanges in reverse order)
        const int dst = trans ? pos : tid;
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
        __syncthreads();
    }
}

template <typename TL, typename TB, bool ACCUM>
rocblas_status getrs_fused_launch(hipStream_t stream, bool trans, int n, int nrhs, Batch<TL> A,
                                  int lda, const int* ipiv, int64_t strideP, Batch<TB> B, int ldb,
                                  Batch<TB> X, int ldx, int bc, const int* skip)
{
    if(bc <= 0 || n <= 0 || nrhs <= 0)
        return rocblas_status_success;
    auto go = [&](auto nmax, auto nr) {
        constexpr int NMAX = decltype(nmax)::value;
        constexpr int NR = decltype(nr)::value;
        hipLaunchKernelGGL((getrs_fused_kernel<TL, TB, NMAX, NR, ACCUM>), dim3(bc), dim3(NMAX), 0,
                           stream, trans, n, nrhs, A, lda, ipiv, strideP, B, ldb, X, ldx, skip);
    };
    auto by_nrhs = [&](auto nmax) {
        if(nrhs <= 1)
            go(nmax, std::integral_constant<int, 1>{});
        else if(nrhs <= 4)
            go(nmax, std::integral_constant<int, 4>{});
        else
            go(nmax, std::integral_constant<int, 8>{});
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
