// trsv_rows.hpp -- triangular solves with the LU factors of getrf for one
// system per workgroup, thread i owning row i of NR right-hand sides in
// registers (n <= NMAX threads).
//
// Each triangle is walked in 32-row diagonal blocks (two barriers per block):
//   * every thread whose row meets the block column loads its 32 entries of
//     that block column into registers right after block k-1's update, so
//     the loads are in flight across block k's two barriers and diagonal
//     solve (one register array of 32 entries; loading a whole block ahead
//     needs two, which halves the resident waves of the FP64 solves);
//   * the half wave owning the diagonal block publishes its right-hand sides
//     and its rows of the block to LDS; every half wave of the workgroup then
//     solves one right-hand side of the block (a 32-step chain in which the
//     value of row k reaches the other lanes through v_readlane);
//   * after the second barrier every row outside the block subtracts its
//     32-entry segment times the block's solution.
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
// rows load nothing and get zeros.  For op N the whole byte offset is the
// per-lane buffer offset, so a column c >= n lies beyond the matrix and the
// hardware returns 0 without a memory access (no branch, no select, and no
// scalar register per column); an inactive lane uses an offset beyond the
// buffer for every column.
template <typename TL, bool TRANS>
__device__ __forceinline__ void load_segment(const MatrixBuffer<TL>& A, int i, int c0, int n,
                                             bool active, TL (&v)[GBS])
{
    if(TRANS)
    {
        const uint32_t vrow = active ? A.col_offset(i) : MatrixBuffer<TL>::OUT_OF_RANGE;
#pragma unroll
        for(int k = 0; k < GBS; ++k)
        {
            const int c = (c0 + k < n) ? c0 + k : n - 1; // uniform
            const TL x = A.load(vrow, A.row_offset(c));
            v[k] = (c0 + k < n) ? x : TL(0);
        }
    }
    else
    {
        // running offset (one register): an inactive lane stays beyond the
        // buffer.  The empty asm hides the increment's value from the
        // optimiser, which would otherwise keep all 32 column offsets live in
        // registers across the whole solve.
        uint32_t off = active ? A.row_offset(i) + A.col_offset(c0) : MatrixBuffer<TL>::OUT_OF_RANGE;
        const uint32_t step = active ? A.ld_bytes : 0u;
#pragma unroll
        for(int k = 0; k < GBS; ++k)
        {
            v[k] = A.load(off, 0);
            off += step;
            opaque(off);
        }
    }
}

// Solve of one 32-row diagonal block, spread over the workgroup: the
// owning half wave has published the block's right-hand sides (sx[i][q],
// rows r0+i) and its rows of the block (sL[i][k] = op(A)[r0+i, r0+k]); half
// wave h of the workgroup solves right-hand sides q = h, h + H, ... (H half
// waves), lane i holding row r0+i, the value of row k reaching the other 31
// lanes through v_readlane (both halves of a wavefront read, each keeps its
// own).  LOWER: forward over k (unit diagonal unless NONUNIT); else backward.
// The block's solution replaces sx.  No barrier inside; the caller brackets
// it with two.
template <typename TL, int NR, bool LOWER, bool NONUNIT>
__device__ __forceinline__ void solve_diag_block(int r0, int n, TL (*sx)[NR],
                                                 const TL (*sL)[GBS + 1])
{
    const int tid = fresh_tid();
    const int i = tid & (GBS - 1);
    const int h = tid >> 5;
    const int H = blockDim.x >> 5;
    for(int q = h; q < NR; q += H) // q differs between the two halves of a wave
    {
        TL y = sx[i][q];
#pragma unroll
        for(int kk = 0; kk < GBS; ++kk)
        {
            const int k = LOWER ? kk : GBS - 1 - kk;
            if(r0 + k < n) // uniform
            {
                const TL lik = sL[i][k];
                if(NONUNIT && i == k)
                    y = y / lik;
                const TL y_lo = readlane_t(y, k);
                const TL y_hi = readlane_t(y, GBS + k);
                const TL yk = (tid & GBS) ? y_hi : y_lo;
                if(LOWER ? (i > k) : (i < k))
                    y = fma_t(-lik, yk, y);
            }
        }
        sx[i][q] = y;
    }
}

// The same step when the diagonal blocks of the factors have been replaced
// by their inverses (invert_diag_blocks: the strictly lower part of a block
// holds the inverse of its unit-lower L block, the upper part with the
// diagonal the inverse of its U block): the block's solution is a product,
// y_i = x_i + sum_{k<i} Linv[i][k] x_k (LOWER) or y_i = sum_{k>=i} Uinv[i][k]
// x_k, so the 32 terms are independent instead of a 32-step chain.  Half
// wave h computes right-hand sides q = h, h + H, ...; its lanes have read all
// of column q before any lane overwrites it (one instruction stream, LDS
// operations in issue order).  Rows and columns past n hold zeros.
template <typename TL, int NR, bool LOWER>
__device__ __forceinline__ void apply_inv_block(TL (*sx)[NR], const TL (*sL)[GBS + 1])
{
    const int tid = fresh_tid();
    const int i = tid & (GBS - 1);
    const int h = tid >> 5;
    const int H = blockDim.x >> 5;
    for(int q = h; q < NR; q += H)
    {
        TL y = LOWER ? sx[i][q] : TL(0);
#pragma unroll 8
        for(int k = 0; k < GBS; ++k)
        {
            // predicated rather than multiplied by zero: an infinite x_k
            // outside the triangle must not turn y into NaN
            const TL t = fma_t(sL[i][k], sx[k][q], y);
            y = (LOWER ? k < i : k >= i) ? t : y;
        }
        sx[i][q] = y;
    }
}

// Forward substitution with the lower triangle of op(A): unit diagonal for
// op N (L), non-unit for op T (U^T); DIAG = 1 forces a non-unit diagonal
// (Cholesky factors).  sx: LDS [2][GBS][NR]; sL: LDS [GBS][GBS+1].  INV
// (op N only): the diagonal blocks hold their inverses.
template <typename TL, int NR, bool TRANS, bool INV = false, int DIAG = -1>
__device__ __forceinline__ void solve_lower_rows(int n, const MatrixBuffer<TL>& A, TL (&x)[NR],
                                                 TL (*sx)[GBS][NR], TL (*sL)[GBS + 1])
{
    static_assert(!(INV && TRANS), "inverted diagonal blocks: op N only");
    const int tid = fresh_tid();
    const int lane32 = tid & (GBS - 1);
    const int nblk = (n + GBS - 1) / GBS;
    TL cur[GBS];
    load_segment<TL, TRANS>(A, tid, 0, n, tid < n, cur);
    for(int blk = 0; blk < nblk; ++blk)
    {
        const int r0 = blk * GBS;
        const int par = blk & 1;
        const bool owner = (tid >> 5) == blk;
        if(owner)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                sx[par][lane32][q] = x[q];
#pragma unroll
            for(int k = 0; k < GBS; ++k)
                sL[lane32][k] = cur[k];
        }
        __syncthreads();
        if constexpr(INV)
            apply_inv_block<TL, NR, true>(sx[par], sL);
        else
            solve_diag_block<TL, NR, true, (DIAG < 0 ? TRANS : DIAG == 1)>(r0, n, sx[par], sL);
        __syncthreads();
        if(owner)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                x[q] = sx[par][lane32][q];
        }
        else if(tid >= r0 + GBS && tid < n)
        {
#pragma unroll
            for(int k = 0; k < GBS; ++k)
                if(r0 + k < n) // uniform
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-cur[k], sx[par][k][q], x[q]);
        }
        if(blk + 1 < nblk) // rows at or below the next block meet its block column
            load_segment<TL, TRANS>(A, tid, r0 + GBS, n, tid >= r0 + GBS && tid < n, cur);
    }
}

// Backward substitution with the upper triangle of op(A): non-unit for op N
// (U), unit diagonal for op T (L^T); DIAG = 1 forces a non-unit diagonal.
// INV as for solve_lower_rows.
template <typename TL, int NR, bool TRANS, bool INV = false, int DIAG = -1>
__device__ __forceinline__ void solve_upper_rows(int n, const MatrixBuffer<TL>& A, TL (&x)[NR],
                                                 TL (*sx)[GBS][NR], TL (*sL)[GBS + 1])
{
    static_assert(!(INV && TRANS), "inverted diagonal blocks: op N only");
    const int tid = fresh_tid();
    const int lane32 = tid & (GBS - 1);
    const int nblk = (n + GBS - 1) / GBS;
    TL cur[GBS];
    {
        const int r0 = (nblk - 1) * GBS;
        load_segment<TL, TRANS>(A, tid, r0, n, tid < n, cur);
    }
    for(int blk = nblk - 1; blk >= 0; --blk)
    {
        const int r0 = blk * GBS;
        const int par = blk & 1;
        const bool owner = (tid >> 5) == blk;
        if(owner)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                sx[par][lane32][q] = x[q];
#pragma unroll
            for(int k = 0; k < GBS; ++k)
                sL[lane32][k] = cur[k];
        }
        __syncthreads();
        if constexpr(INV)
            apply_inv_block<TL, NR, false>(sx[par], sL);
        else
            solve_diag_block<TL, NR, false, (DIAG < 0 ? !TRANS : DIAG == 1)>(r0, n, sx[par], sL);
        __syncthreads();
        if(owner)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                x[q] = sx[par][lane32][q];
        }
        else if(tid < r0)
        {
#pragma unroll
            for(int k = GBS - 1; k >= 0; --k)
                if(r0 + k < n) // uniform
#pragma unroll
                    for(int q = 0; q < NR; ++q)
                        x[q] = fma_t(-cur[k], sx[par][k][q], x[q]);
        }
        if(blk > 0) // rows above this block meet the previous block column
            load_segment<TL, TRANS>(A, tid, r0 - GBS, n, tid < r0, cur);
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
    const int tid = fresh_tid();
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
