// lu_blocked.hip -- right-looking blocked LU and blocked triangular solves for
// any m, n (the path for matrices beyond the register-resident kernels).
//
// getrf, per block column j of width jb (reference structure:
// library/src/lapack/roclapack_getrf.hpp:608-752, without rocBLAS):
//   1. panel_kernel: one workgroup per matrix factors A[j:m, j:j+jb] with
//      partial pivoting (workgroup arg-max per column), interchanging rows
//      inside the panel at each column; then applies the panel's interchanges
//      to every other column at once (roclapack_getf2.hpp:41-99 semantics:
//      the same final rows); u12_kernel then solves U12 = L11^-1 A12 for the
//      columns right of the panel, one thread per column.
//   2. gemm_sub_kernel: A22 -= L21 * U12 on MFMA (v_mfma_f64_16x16x4f64 /
//      v_mfma_f32_16x16x4f32), 64x64 tiles.
// getrs (roclapack_getrs.hpp:117-150): laswp, then for each 64-row diagonal
// block a substitution kernel followed by an MFMA GEMM update of the rest.
#include "internal.hpp"
#include "mfma.hpp"

#include <climits>

namespace rsa
{

constexpr int PANEL_NB = 32; // panel width
constexpr int PANEL_THREADS = 256;
constexpr int PANEL_TALL = 2048; // panels of this many rows or more: 2 * PANEL_THREADS
constexpr int TRSM_NB = 64; // getrs diagonal block
constexpr int64_t MAX_GRID_YZ = 65535; // grid y/z limit: larger batches are launched in chunks

// ------------------------------------------------------------------ info
__global__ void reset_info_kernel(int* info, int bc)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < bc)
        info[i] = 0;
}

rocblas_status reset_info(hipStream_t stream, int* info, int bc)
{
    if(bc <= 0)
        return rocblas_status_success;
    hipLaunchKernelGGL(reset_info_kernel, dim3(ceil_div(bc, 256)), dim3(256), 0, stream, info, bc);
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

// ------------------------------------------------------------ panel + trsm
template <typename T, bool PIVOT, int NT>
__global__ __launch_bounds__(NT) void panel_kernel(int m, int n, Batch<T> AA, int lda,
                                                              int* __restrict__ ipivA,
                                                              int64_t strideP,
                                                              int* __restrict__ info, int j0,
                                                              int jb)
{
    constexpr int NW = NT / 64;
    __shared__ T s_key[NW];
    __shared__ int s_idx[NW];
    __shared__ T s_val[NW]; // each wavefront's candidate value
    __shared__ int s_swap[PANEL_NB]; // row interchanged with row j0 + k (itself: none)
    __shared__ int s_pos[2 * PANEL_NB]; // rows touched by the panel's interchanges
    __shared__ int s_src[2 * PANEL_NB]; // original row whose entries end at s_pos[q]
    __shared__ int s_cnt;

    const int64_t b = blockIdx.x;
    T* A = AA[b];
    const int tid = threadIdx.x;
    const int wid = tid / 64;
    const int lane = tid % 64;
    auto at = [&](int i, int j) -> T& { return A[i + int64_t(j) * lda]; };

    for(int k = 0; k < jb; ++k)
    {
        const int c = j0 + k;
        int p = c;
        T pv;
        if(PIVOT)
        {
            T key = T(-1), val = T(0);
            int idx = INT_MAX;
#pragma unroll 4
            for(int i = c + tid; i < m; i += NT)
            {
                const T a = at(i, c);
                T kk = pivot_key(a, i == c);
                if(kk > key)
                {
                    key = kk;
                    idx = i;
                    val = a;
                }
            }
            const int my_idx = idx;
            wave_argmax(key, idx);
            if(lane == 0)
            {
                s_key[wid] = key;
                s_idx[wid] = idx;
            }
            // the lane holding the winning row publishes its value (rows are
            // distinct across lanes), so no global read of it follows
            if(my_idx == idx && idx != INT_MAX)
                s_val[wid] = val;
            __syncthreads();
            // every thread combines the wavefronts' candidates itself (no
            // serial section and no second barrier)
            T bk = s_key[0];
            int bw = 0;
            p = s_idx[0];
            for(int w = 1; w < NW; ++w)
                if(s_key[w] > bk || (s_key[w] == bk && s_idx[w] < p))
                {
                    bk = s_key[w];
                    p = s_idx[w];
                    bw = w;
                }
            pv = s_val[bw];
            if(tid == 0)
            {
                ipivA[b * strideP + c] = p + 1;
                if(pv == T(0) && info[b] == 0)
                    info[b] = c + 1;
            }
        }
        else
        {
            pv = at(c, c);
            if(tid == 0 && pv == T(0) && info[b] == 0)
                info[b] = c + 1;
        }
        const bool nz = pv != T(0);
        if(tid == 0)
            s_swap[k] = (nz && p != c) ? p : c;
        // interchange rows c and p inside the panel (the other columns take
        // all of the panel's interchanges at once after it)
        const int jend = j0 + jb;
        if(nz && p != c)
        {
            for(int j = j0 + tid; j < jend; j += NT)
            {
                T t = at(c, j);
                at(c, j) = at(p, j);
                at(p, j) = t;
            }
        }
        __syncthreads();
        // scale the column and update the rest of the panel.  A zero pivot
        // leaves the column alone with pivoting (every candidate is zero) and
        // multiplies it by zero without (the reference npvt kernels keep the
        // scale at 0: roclapack_getf2_specialized_kernels.hpp:185-188,427-430)
        const T rcp = nz ? T(1) / pv : T(0);
        // the pivot row's panel entries right of c, then per row all loads
        // before any store, so they are in flight together
        T u[PANEL_NB];
#pragma unroll
        for(int q = 0; q < PANEL_NB; ++q)
        {
            const int j = j0 + q;
            u[q] = (j > c && j < jend) ? at(c, j) : T(0);
        }
        // two rows per pass (i and i + NT): both rows' loads are
        // issued before either row's stores
        for(int i = c + 1 + tid; i < m; i += 2 * NT)
        {
            const int i1 = i + NT;
            const bool has1 = i1 < m;
            T l0 = at(i, c);
            T l1 = has1 ? at(i1, c) : T(0);
            T v0[PANEL_NB], v1[PANEL_NB];
#pragma unroll
            for(int q = 0; q < PANEL_NB; ++q)
            {
                const int j = j0 + q;
                const bool in = j > c && j < jend;
                v0[q] = in ? at(i, j) : T(0);
                v1[q] = (in && has1) ? at(i1, j) : T(0);
            }
            if(nz || !PIVOT)
            {
                l0 *= rcp;
                l1 *= rcp;
                at(i, c) = l0;
                if(has1)
                    at(i1, c) = l1;
            }
#pragma unroll
            for(int q = 0; q < PANEL_NB; ++q)
            {
                const int j = j0 + q;
                if(j > c && j < jend)
                {
                    at(i, j) = fma_t(-l0, u[q], v0[q]);
                    if(has1)
                        at(i1, j) = fma_t(-l1, u[q], v1[q]);
                }
            }
        }
        __syncthreads();
    }

    // the panel's interchanges on every column outside it: composed into one
    // map of at most 2 jb rows (s_src[q]: the original row whose entries end
    // in row s_pos[q]), then each column reads those rows and writes them
    // back permuted
    if(PIVOT && n > jb)
    {
        if(wid == 0)
        {
            // lane q keeps slot q (row pos, source src); ballots find a row's
            // slot, a new row takes the next free one
            int pos = -1, src = -1, cnt = 0;
            for(int k = 0; k < jb; ++k)
            {
                const int r = j0 + k, r2 = s_swap[k];
                if(r2 == r) // uniform
                    continue;
                int qs[2];
#pragma unroll
                for(int h = 0; h < 2; ++h)
                {
                    const int row = h ? r2 : r;
                    const uint64_t hit = __builtin_amdgcn_ballot_w64(pos == row);
                    int q = hit ? __builtin_ffsll(hit) - 1 : cnt;
                    if(!hit)
                    {
                        if(lane == q)
                        {
                            pos = row;
                            src = row;
                        }
                        ++cnt;
                    }
                    qs[h] = q;
                }
                const int sa = __builtin_amdgcn_readlane(src, qs[0]);
                const int sb = __builtin_amdgcn_readlane(src, qs[1]);
                if(lane == qs[0])
                    src = sb;
                if(lane == qs[1])
                    src = sa;
            }
            if(lane < cnt)
            {
                s_pos[lane] = pos;
                s_src[lane] = src;
            }
            if(lane == 0)
                s_cnt = cnt;
        }
        __syncthreads();
        const int cnt = s_cnt;
        const int nout = n - jb; // columns left and right of the panel
        for(int e = tid; e < nout; e += NT)
        {
            const int j = e < j0 ? e : e + jb;
            T v[2 * PANEL_NB];
#pragma unroll
            for(int q = 0; q < 2 * PANEL_NB; ++q)
                v[q] = q < cnt ? at(s_src[q], j) : T(0);
#pragma unroll
            for(int q = 0; q < 2 * PANEL_NB; ++q)
                if(q < cnt)
                    at(s_pos[q], j) = v[q];
        }
        __syncthreads();
    }
}

// U12 = L11^-1 A12 for the columns right of a finished panel (L11 = the unit
// lower triangle of A[j0:j0+jb, j0:j0+jb]): one thread per column, 64
// columns per workgroup, L11 staged in LDS (zero outside jb x jb).
template <typename T>
__global__ __launch_bounds__(64) void u12_kernel(int n, Batch<T> AA, int lda, int j0, int jb,
                                                 int64_t boff)
{
    __shared__ T sL[PANEL_NB * PANEL_NB];
    const int64_t b = blockIdx.y + boff;
    T* A = AA[b];
    const int tid = threadIdx.x;
    for(int e = tid; e < PANEL_NB * PANEL_NB; e += 64)
    {
        const int i = e % PANEL_NB, k = e / PANEL_NB;
        sL[e] = (i < jb && k < jb) ? A[j0 + i + int64_t(j0 + k) * lda] : T(0);
    }
    __syncthreads();
    const int j = j0 + jb + blockIdx.x * 64 + tid;
    if(j >= n)
        return;
    T* col = A + int64_t(j) * lda + j0;
    T x[PANEL_NB];
#pragma unroll
    for(int i = 0; i < PANEL_NB; ++i)
        x[i] = i < jb ? col[i] : T(0);
    // L11 read through an offset that the optimiser must take as computed
    // from x[k] every two columns (an empty asm tied to x[k]), so its 496
    // entries are not all loaded ahead of the substitution (and spilled)
    const T* L = sL;
#pragma unroll
    for(int k = 0; k < PANEL_NB; ++k)
    {
        if((k & 1) == 0)
        {
            uint32_t lo = 0;
            asm volatile("" : "+v"(lo) : "v"(x[k]));
            L = sL + lo;
        }
#pragma unroll
        for(int i = k + 1; i < PANEL_NB; ++i)
            x[i] = fma_t(-L[i + k * PANEL_NB], x[k], x[i]);
    }
#pragma unroll
    for(int i = 0; i < PANEL_NB; ++i)
        if(i < jb)
            col[i] = x[i];
}

// ------------------------------------------------------------------ GEMM
// C[M x N] -= op(A)[M x K] * B[K x N], column-major, one 64x64 tile per
// workgroup (4 waves, each 32x32 = 2x2 MFMA 16x16 blocks), K staged through
// LDS in chunks of 16 (MFMA operand/result layouts: mfma.hpp).
constexpr int GT = 64; // tile
constexpr int GK = 16; // k chunk

template <typename T, bool TRANSA>
__global__ __launch_bounds__(256) void gemm_sub_kernel(int M, int N, int K, Batch<T> AA,
                                                       int64_t offA, int lda, Batch<T> BB,
                                                       int64_t offB, int ldb, Batch<T> CC,
                                                       int64_t offC, int ldc,
                                                       const int* __restrict__ skip, int64_t boff)
{
    using acc_t = typename mfma16<T>::acc_t;
    __shared__ T sA[GK][GT]; // sA[k][i]
    __shared__ T sB[GK][GT]; // sB[k][j]

    const int64_t b = blockIdx.z + boff;
    if(skip && skip[b] != 0)
        return;
    const T* A = AA[b] + offA;
    const T* Bm = BB[b] + offB;
    T* C = CC[b] + offC;
    const int i0 = blockIdx.x * GT;
    const int j0 = blockIdx.y * GT;
    const int tid = threadIdx.x;
    const int wid = tid / 64;
    const int lane = tid % 64;
    const int wi = (wid % 2) * 32;
    const int wj = (wid / 2) * 32;

    acc_t acc[2][2];
#pragma unroll
    for(int s = 0; s < 2; ++s)
#pragma unroll
        for(int t = 0; t < 2; ++t)
            acc[s][t] = acc_t{0, 0, 0, 0};

    for(int k0 = 0; k0 < K; k0 += GK)
    {
        // stage A: GT x GK, B: GK x GT (zero padded)
#pragma unroll
        for(int e = tid; e < GT * GK; e += 256)
        {
            int i, k;
            T v;
            if(!TRANSA)
            {
                i = e % GT;
                k = e / GT;
                v = (i0 + i < M && k0 + k < K) ? A[(i0 + i) + int64_t(k0 + k) * lda] : T(0);
            }
            else
            {
                k = e % GK;
                i = e / GK;
                v = (i0 + i < M && k0 + k < K) ? A[(k0 + k) + int64_t(i0 + i) * lda] : T(0);
            }
            sA[k][i] = v;
            const int kb = e % GK, jb = e / GK;
            sB[kb][jb] = (k0 + kb < K && j0 + jb < N) ? Bm[(k0 + kb) + int64_t(j0 + jb) * ldb]
                                                      : T(0);
        }
        __syncthreads();
#pragma unroll
        for(int kk = 0; kk < GK; kk += 4)
        {
            T a[2], bv[2];
#pragma unroll
            for(int s = 0; s < 2; ++s)
                a[s] = sA[kk + lane / 16][wi + s * 16 + lane % 16];
#pragma unroll
            for(int t = 0; t < 2; ++t)
                bv[t] = sB[kk + lane / 16][wj + t * 16 + lane % 16];
#pragma unroll
            for(int s = 0; s < 2; ++s)
#pragma unroll
                for(int t = 0; t < 2; ++t)
                    acc[s][t] = mfma16<T>::run(a[s], bv[t], acc[s][t]);
        }
        __syncthreads();
    }

#pragma unroll
    for(int s = 0; s < 2; ++s)
#pragma unroll
        for(int t = 0; t < 2; ++t)
        {
            const int j = j0 + wj + t * 16 + lane % 16;
#pragma unroll
            for(int r = 0; r < 4; ++r)
            {
                const int i = i0 + wi + s * 16 + mfma16<T>::row(lane, r);
                if(i < M && j < N)
                {
                    T& c = C[i + int64_t(j) * ldc];
                    c = c - acc[s][t][r];
                }
            }
        }
}

template <typename T>
static rocblas_status gemm_sub(hipStream_t stream, bool transA, int M, int N, int K, Batch<T> A,
                               int64_t offA, int lda, Batch<T> B, int64_t offB, int ldb,
                               Batch<T> C, int64_t offC, int ldc, int bc, const int* skip)
{
    if(M <= 0 || N <= 0 || K <= 0 || bc <= 0)
        return rocblas_status_success;
    for(int64_t b0 = 0; b0 < bc; b0 += MAX_GRID_YZ)
    {
        const int nb = int((bc - b0) < MAX_GRID_YZ ? (bc - b0) : MAX_GRID_YZ);
        dim3 grid(ceil_div(M, GT), ceil_div(N, GT), nb);
        if(transA)
            hipLaunchKernelGGL((gemm_sub_kernel<T, true>), grid, dim3(256), 0, stream, M, N, K, A,
                               offA, lda, B, offB, ldb, C, offC, ldc, skip, b0);
        else
            hipLaunchKernelGGL((gemm_sub_kernel<T, false>), grid, dim3(256), 0, stream, M, N, K,
                               A, offA, lda, B, offB, ldb, C, offC, ldc, skip, b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

// ----------------------------------------------------------------- getrf
template <typename T>
rocblas_status getrf_blocked(hipStream_t stream, int m, int n, Batch<T> A, int lda, int* ipiv,
                             int64_t strideP, int* info, int bc, bool pivot)
{
    rocblas_status st = reset_info(stream, info, bc);
    if(st != rocblas_status_success)
        return st;
    const int kmax = m < n ? m : n;
    for(int j = 0; j < kmax; j += PANEL_NB)
    {
        const int jb = (kmax - j) < PANEL_NB ? (kmax - j) : PANEL_NB;
        // tall panels (2048 rows or more) with 512 threads: half the rows
        // per thread in each column step; shorter ones with 256, whose
        // update keeps two rows' loads in flight without spilling
        if(pivot && m >= PANEL_TALL)
            hipLaunchKernelGGL((panel_kernel<T, true, 2 * PANEL_THREADS>), dim3(bc),
                               dim3(2 * PANEL_THREADS), 0, stream, m, n, A, lda, ipiv, strideP,
                               info, j, jb);
        else if(pivot)
            hipLaunchKernelGGL((panel_kernel<T, true, PANEL_THREADS>), dim3(bc), dim3(PANEL_THREADS),
                               0, stream, m, n, A, lda, ipiv, strideP, info, j, jb);
        else if(m >= PANEL_TALL)
            hipLaunchKernelGGL((panel_kernel<T, false, 2 * PANEL_THREADS>), dim3(bc),
                               dim3(2 * PANEL_THREADS), 0, stream, m, n, A, lda, ipiv, strideP,
                               info, j, jb);
        else
            hipLaunchKernelGGL((panel_kernel<T, false, PANEL_THREADS>), dim3(bc),
                               dim3(PANEL_THREADS), 0, stream, m, n, A, lda, ipiv, strideP, info,
                               j, jb);
        RSA_LAUNCH_CHECK();
        const int r0 = j + jb;
        if(r0 < n)
            for(int64_t b0 = 0; b0 < bc; b0 += MAX_GRID_YZ)
            {
                const int nb = int((bc - b0) < MAX_GRID_YZ ? (bc - b0) : MAX_GRID_YZ);
                hipLaunchKernelGGL((u12_kernel<T>), dim3(ceil_div(n - r0, 64), nb), dim3(64), 0,
                                   stream, n, A, lda, j, jb, b0);
                RSA_LAUNCH_CHECK();
            }
        if(r0 < m && r0 < n)
        {
            st = gemm_sub<T>(stream, false, m - r0, n - r0, jb, A, r0 + int64_t(j) * lda, lda, A,
                             j + int64_t(r0) * lda, lda, A, r0 + int64_t(r0) * lda, lda, bc,
                             nullptr);
            if(st != rocblas_status_success)
                return st;
        }
    }
    return rocblas_status_success;
}

// ----------------------------------------------------------------- laswp
// One thread per column; the interchanges of one column run in order
// (rocauxiliary_laswp.hpp:41-62).
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(int ncols, Batch<T> AA, int lda, int k1,
                                                    int k2, const int* __restrict__ ipivA,
                                                    int64_t strideP, int incx,
                                                    const int* __restrict__ skip, int64_t boff)
{
    const int64_t b = blockIdx.y + boff;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= ncols || (skip && skip[b] != 0))
        return;
    T* A = AA[b] + int64_t(c) * lda;
    const int* ipiv = ipivA + b * strideP;
    int start, end, inc;
    if(incx < 0)
    {
        start = k2;
        end = k1 - 1;
        inc = -1;
        incx = -incx;
    }
    else
    {
        start = k1;
        end = k2 + 1;
        inc = 1;
    }
    for(int i = start; i != end; i += inc)
    {
        const int ex = ipiv[k1 + (i - k1) * incx - 1];
        // a pivot outside the column's storage is ignored instead of addressed
        if(ex != i && ex >= 1 && ex <= lda)
        {
            T t = A[i - 1];
            A[i - 1] = A[ex - 1];
            A[ex - 1] = t;
        }
    }
}

template <typename T>
rocblas_status laswp(hipStream_t stream, int ncols, Batch<T> A, int lda, int k1, int k2,
                     const int* ipiv, int64_t strideP, int incx, int bc, const int* skip_info)
{
    if(ncols <= 0 || bc <= 0)
        return rocblas_status_success;
    for(int64_t b0 = 0; b0 < bc; b0 += MAX_GRID_YZ)
    {
        const int nb = int((bc - b0) < MAX_GRID_YZ ? (bc - b0) : MAX_GRID_YZ);
        dim3 grid(ceil_div(ncols, 256), nb);
        hipLaunchKernelGGL((laswp_kernel<T>), grid, dim3(256), 0, stream, ncols, A, lda, k1, k2,
                           ipiv, strideP, incx, skip_info, b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

// ------------------------------------------------------------ substitution
// Solves one diagonal block (rows r0..r0+rb) of the LU against all right-hand
// sides, one thread per column of B, block in LDS.  kind: 0 = L unit lower
// (forward), 1 = U non-unit upper (backward), 2 = U^T (forward, non-unit),
// 3 = L^T (backward, unit).  Column-oriented order with true division, as
// the reference substitution kernels.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void trsm_block_kernel(int r0, int rb, int nrhs, Batch<T> AA,
                                                         int lda, Batch<T> BB, int ldb,
                                                         const int* __restrict__ skip,
                                                         int64_t boff)
{
    __shared__ T sD[TRSM_NB * TRSM_NB];
    const int64_t b = blockIdx.y + boff;
    if(skip && skip[b] != 0)
        return;
    const T* A = AA[b];
    T* Bm = BB[b];
    const int tid = threadIdx.x;
    for(int e = tid; e < rb * rb; e += 256)
    {
        const int i = e % rb, k = e / rb;
        sD[i + k * TRSM_NB] = A[(r0 + i) + int64_t(r0 + k) * lda];
    }
    __syncthreads();
    const int c = blockIdx.x * 256 + tid;
    if(c >= nrhs)
        return;
    T x[TRSM_NB];
#pragma unroll
    for(int i = 0; i < TRSM_NB; ++i)
        x[i] = i < rb ? Bm[(r0 + i) + int64_t(c) * ldb] : T(0);
    auto D = [&](int i, int k) { return sD[i + k * TRSM_NB]; };
    if(KIND == 0)
    {
        // the block read through an offset tied to x[k] every two columns
        // (an empty asm), so its entries are not all loaded ahead and spilled
        const T* Dp = sD;
#pragma unroll
        for(int k = 0; k < TRSM_NB; ++k)
        {
            if((k & 1) == 0)
            {
                uint32_t lo = 0;
                asm volatile("" : "+v"(lo) : "v"(x[k]));
                Dp = sD + lo;
            }
#pragma unroll
            for(int i = k + 1; i < TRSM_NB; ++i)
                if(i < rb)
                    x[i] = fma_t(-Dp[i + k * TRSM_NB], x[k], x[i]);
        }
    }
    else if(KIND == 1)
    {
#pragma unroll
        for(int k = TRSM_NB - 1; k >= 0; --k)
        {
            if(k < rb)
            {
                x[k] = x[k] / D(k, k);
#pragma unroll
                for(int i = 0; i < k; ++i)
                    x[i] = fma_t(-D(i, k), x[k], x[i]);
            }
        }
    }
    else if(KIND == 2)
    {
#pragma unroll
        for(int k = 0; k < TRSM_NB; ++k)
        {
            if(k < rb)
            {
                x[k] = x[k] / D(k, k);
#pragma unroll
                for(int i = k + 1; i < TRSM_NB; ++i)
                    if(i < rb)
                        x[i] = fma_t(-D(k, i), x[k], x[i]);
            }
        }
    }
    else
    {
#pragma unroll
        for(int k = TRSM_NB - 1; k > 0; --k)
        {
            if(k < rb)
            {
#pragma unroll
                for(int i = 0; i < k; ++i)
                    x[i] = fma_t(-D(k, i), x[k], x[i]);
            }
        }
    }
#pragma unroll
    for(int i = 0; i < TRSM_NB; ++i)
        if(i < rb)
            Bm[(r0 + i) + int64_t(c) * ldb] = x[i];
}

template <typename T, int KIND>
static rocblas_status trsm_block(hipStream_t stream, int r0, int rb, int nrhs, Batch<T> A, int lda,
                                 Batch<T> B, int ldb, int bc, const int* skip)
{
    for(int64_t b0 = 0; b0 < bc; b0 += MAX_GRID_YZ)
    {
        const int nb = int((bc - b0) < MAX_GRID_YZ ? (bc - b0) : MAX_GRID_YZ);
        dim3 grid(ceil_div(nrhs, 256), nb);
        hipLaunchKernelGGL((trsm_block_kernel<T, KIND>), grid, dim3(256), 0, stream, r0, rb, nrhs,
                           A, lda, B, ldb, skip, b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

template <typename T>
rocblas_status getrs_blocked(hipStream_t stream, bool trans, int n, int nrhs, Batch<T> A, int lda,
                             const int* ipiv, int64_t strideP, Batch<T> B, int ldb, int bc,
                             bool pivot, const int* skip)
{
    rocblas_status st = rocblas_status_success;
#define RSA_TRY(x)                            \
    do                                        \
    {                                         \
        st = (x);                             \
        if(st != rocblas_status_success)      \
            return st;                        \
    } while(0)
    const int nb = TRSM_NB;
    if(!trans)
    {
        if(pivot)
            RSA_TRY(laswp<T>(stream, nrhs, B, ldb, 1, n, ipiv, strideP, 1, bc, skip));
        for(int r0 = 0; r0 < n; r0 += nb)
        {
            const int rb = (n - r0) < nb ? (n - r0) : nb;
            RSA_TRY((trsm_block<T, 0>(stream, r0, rb, nrhs, A, lda, B, ldb, bc, skip)));
            const int r1 = r0 + rb;
            RSA_TRY(gemm_sub<T>(stream, false, n - r1, nrhs, rb, A, r1 + int64_t(r0) * lda, lda, B,
                                r0, ldb, B, r1, ldb, bc, skip));
        }
        const int last = ((n - 1) / nb) * nb;
        for(int r0 = last; r0 >= 0; r0 -= nb)
        {
            const int rb = (n - r0) < nb ? (n - r0) : nb;
            RSA_TRY((trsm_block<T, 1>(stream, r0, rb, nrhs, A, lda, B, ldb, bc, skip)));
            RSA_TRY(gemm_sub<T>(stream, false, r0, nrhs, rb, A, int64_t(r0) * lda, lda, B, r0,
                                ldb, B, 0, ldb, bc, skip));
        }
    }
    else
    {
        for(int r0 = 0; r0 < n; r0 += nb)
        {
            const int rb = (n - r0) < nb ? (n - r0) : nb;
            RSA_TRY((trsm_block<T, 2>(stream, r0, rb, nrhs, A, lda, B, ldb, bc, skip)));
            const int r1 = r0 + rb;
            // B[r1:n] -= (U[r0:r1, r1:n])^T B[r0:r1]
            RSA_TRY(gemm_sub<T>(stream, true, n - r1, nrhs, rb, A, r0 + int64_t(r1) * lda, lda, B,
                                r0, ldb, B, r1, ldb, bc, skip));
        }
        const int last = ((n - 1) / nb) * nb;
        for(int r0 = last; r0 >= 0; r0 -= nb)
        {
            const int rb = (n - r0) < nb ? (n - r0) : nb;
            RSA_TRY((trsm_block<T, 3>(stream, r0, rb, nrhs, A, lda, B, ldb, bc, skip)));
            // B[0:r0] -= (L[r0:r1, 0:r0])^T B[r0:r1]
            RSA_TRY(gemm_sub<T>(stream, true, r0, nrhs, rb, A, r0, lda, B, r0, ldb, B, 0, ldb, bc,
                                skip));
        }
        if(pivot)
            RSA_TRY(laswp<T>(stream, nrhs, B, ldb, 1, n, ipiv, strideP, -1, bc, skip));
    }
#undef RSA_TRY
    return rocblas_status_success;
}

#define RSA_INST(T)                                                                              \
    template rocblas_status getrf_blocked<T>(hipStream_t, int, int, Batch<T>, int, int*, int64_t, \
                                             int*, int, bool);                                   \
    template rocblas_status getrs_blocked<T>(hipStream_t, bool, int, int, Batch<T>, int,          \
                                             const int*, int64_t, Batch<T>, int, int, bool,       \
                                             const int*);                                        \
    template rocblas_status laswp<T>(hipStream_t, int, Batch<T>, int, int, int, const int*,       \
                                     int64_t, int, int, const int*);
RSA_INST(float)
RSA_INST(double)
#undef RSA_INST

} // namespace rsa
