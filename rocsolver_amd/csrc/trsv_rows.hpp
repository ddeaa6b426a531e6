// trsv_rows.hpp -- triangular solves with the LU factors of getrf for one
// system per workgroup, thread i owning row i of NR right-hand sides in
// registers (n <= NMAX threads).
//
// Each triangle is walked in 32-row diagonal blocks (one barrier per block):
//   * every thread whose row meets the block column loads its 32 entries of
//     that block column into registers one block ahead of use, so the loads
//     of block k+1 are in flight while block k is solved;
//   * the half wave owning the diagonal block solves it in registers: the
//     value of row k reaches the other 31 lanes through v_readlane (no LDS,
//     no barrier), and publishes the block's solution to LDS;
//   * after the barrier every row outside the block subtracts its 32-entry
//     segment times the block's solution.
// The arithmetic is the column-oriented substitution of the reference
// (library/src/lapack/roclapack_getrs.hpp:117-150 applies laswp, then the
// unit-lower L, then the non-unit upper U for op = N; U^T, L^T, then the
// reverse interchanges for op = T): for every row the updates are applied
// in increasing (forward) or decreasing (backward) column order with a
// fused multiply-add, and the diagonal is applied by true division.
#pragma once

#include "internal.hpp"

namespace rsa
{

constexpr int GBS = 32; // diagonal block

// Entries c0 .. c0+31 of row i of op(A) restricted to one block column:
// op N reads A[i, c] (coalesced across threads), op T reads A[c, i] (32
// consecutive entries per thread).  Entries with c >= n are zero; inactive
// rows load nothing (out-of-range buffer offset) and get zeros.
template <typename TL, bool TRANS>
__device__ __forceinline__ void load_segment(const MatrixBuffer<TL>& A, int i, int c0, int n,
                                             bool active, TL (&v)[GBS])
{
    const uint32_t vrow = TRANS ? A.col_offset(i) : A.row_offset(i);
    const uint32_t voff = active ? vrow : MatrixBuffer<TL>::OUT_OF_RANGE;
#pragma unroll
    for(int k = 0; k < GBS; ++k)
    {
        const int c = (c0 + k < n) ? c0 + k : n - 1; // uniform
        const uint32_t soff = TRANS ? A.row_offset(c) : A.col_offset(c);
        const TL x = A.load(voff, soff);
        v[k] = (c0 + k < n) ? x : TL(0);
    }
}

// Forward substitution with the lower triangle of op(A): unit diagonal for
// op N (L), non-unit for op T (U^T).  sx: LDS [2][GBS][NR].
template <typename TL, int NR, bool TRANS>
__device__ __forceinline__ void solve_lower_rows(int n, const MatrixBuffer<TL>& A, TL (&x)[NR],
                                                 TL (*sx)[GBS][NR])
{
    const int tid = threadIdx.x;
    const int lane32 = tid & (GBS - 1);
    const int nblk = (n + GBS - 1) / GBS;
    TL cur[GBS], nxt[GBS];
    load_segment<TL, TRANS>(A, tid, 0, n, tid < n, cur);
    for(int blk = 0; blk < nblk; ++blk)
    {
        const int r0 = blk * GBS;
        const int par = blk & 1;
        if(blk + 1 < nblk) // rows at or below the next block meet its block column
            load_segment<TL, TRANS>(A, tid, r0 + GBS, n, tid >= r0 + GBS && tid < n, nxt);
        if((tid >> 5) == blk)
        {
            const int base = r0 & 63; // first lane of the owning half wave
#pragma unroll
            for(int k = 0; k < GBS; ++k)
            {
                if(r0 + k < n) // uniform
                {
                    if(TRANS && lane32 == k)
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = x[q] / cur[k];
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                    {
                        const TL xk = readlane_t(x[q], base + k);
                        if(lane32 > k)
                            x[q] = fma_t(-cur[k], xk, x[q]);
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
#pragma unroll
            for(int k = 0; k < GBS; ++k)
                if(r0 + k < n) // uniform
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-cur[k], sx[par][k][q], x[q]);
        }
#pragma unroll
        for(int k = 0; k < GBS; ++k)
            cur[k] = nxt[k];
    }
}

// Backward substitution with the upper triangle of op(A): non-unit for op N
// (U), unit diagonal for op T (L^T).
template <typename TL, int NR, bool TRANS>
__device__ __forceinline__ void solve_upper_rows(int n, const MatrixBuffer<TL>& A, TL (&x)[NR],
                                                 TL (*sx)[GBS][NR])
{
    const int tid = threadIdx.x;
    const int lane32 = tid & (GBS - 1);
    const int nblk = (n + GBS - 1) / GBS;
    TL cur[GBS], nxt[GBS];
    {
        const int r0 = (nblk - 1) * GBS;
        load_segment<TL, TRANS>(A, tid, r0, n, tid < n, cur);
    }
    for(int blk = nblk - 1; blk >= 0; --blk)
    {
        const int r0 = blk * GBS;
        const int par = blk & 1;
        if(blk > 0) // rows above this block meet the previous block column
            load_segment<TL, TRANS>(A, tid, r0 - GBS, n, tid < r0, nxt);
        if((tid >> 5) == blk)
        {
            const int base = r0 & 63;
#pragma unroll
            for(int k = GBS - 1; k >= 0; --k)
            {
                if(r0 + k < n) // uniform
                {
                    if(!TRANS && lane32 == k)
#pragma unroll
                        for(int q = 0; q < NR; ++q)
                            x[q] = x[q] / cur[k];
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                    {
                        const TL xk = readlane_t(x[q], base + k);
                        if(lane32 < k)
                            x[q] = fma_t(-cur[k], xk, x[q]);
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
#pragma unroll
            for(int k = GBS - 1; k >= 0; --k)
                if(r0 + k < n) // uniform
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-cur[k], sx[par][k][q], x[q]);
        }
#pragma unroll
        for(int k = 0; k < GBS; ++k)
            cur[k] = nxt[k];
    }
}

// Interchanges of ipiv (1-based, LAPACK order) as a map between rows.
// Returns the final position of original row tid after applying the
// interchanges forward (op N) or in reverse (op T), and for op N fills
// ssrc[pos] = original row landing on pos.  A pivot outside [1, n] is
// ignored.  Two barriers; ssrc and sipiv are LDS arrays of NMAX ints.
__device__ __forceinline__ int pivot_position(int n, const int* __restrict__ ipiv, bool trans,
                                              int* sipiv, int* ssrc)
{
    const int tid = threadIdx.x;
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
        ssrc[pos] = tid;
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
    return pos;
}

} // namespace rsa
