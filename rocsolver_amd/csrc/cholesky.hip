// cholesky.hip -- batched Cholesky factorisation (potf2 / potrf), solve with
// the factor (potrs) and factor-and-solve (posv), uplo lower (A = L L^T) or
// upper (A = U^T U).  Only the referenced triangle is read or written.
//
// Factorisation: one workgroup of 256 threads per matrix, left-looking over
// 32-column block columns, the rows of the block column spread one per
// thread (chunks of 256 rows when n is larger):
//   1. every row of the block column subtracts its products with the finished
//      columns, 32 at a time (the 32 rows of the diagonal block staged in LDS);
//   2. the wavefront holding the diagonal block factors it in registers
//      (right-looking; entries of column k reach the other lanes through
//      v_readlane);
//   3. every row below solves its 32 entries against the factored block
//      (multiplying by the stored reciprocals), and the block column is
//      written.
// Per entry the order of operations is that of the reference's unblocked
// potf2 (library/src/lapack/roclapack_potf2.hpp:160-250): t = a_jj minus the
// squares of row j's finished entries (one fma each, k increasing); t <= 0 ->
// info = first j + 1, a_jj = t and the column below is scaled by 0, and the
// factorisation goes on (sqrtDiagOnward, :10-37); otherwise a_jj = sqrt(t)
// and the entries below are (a_ij minus their products, k increasing) times
// 1 / a_jj.  So the factor equals the oracle's restatement bit for bit; the
// reference's own dot/gemv reductions differ from it by rounding only.
//
// Solve: for n <= 512 one workgroup per system with the row solvers of the
// LU path (trsv_rows.hpp) on non-unit triangles: L then L^T, or U^T then U
// (library/src/lapack/roclapack_potrs.hpp:121-148); larger orders use one
// thread per (system, right-hand side) substitution.
#include "trsv_rows.hpp"

#include <cmath>

namespace rsa
{

constexpr int CHB = 32; // block-column width
constexpr int CHT = 256; // threads per workgroup

// Element (i, j), i >= j, of the lower factor L: A[i + j lda] with lower
// storage, A[j + i lda] with upper storage (which holds U = L^T).
template <typename T>
struct LowerView
{
    T* a;
    int lda;
    bool up;
    __device__ __forceinline__ T& operator()(int i, int j) const
    {
        return up ? a[j + int64_t(i) * lda] : a[i + int64_t(j) * lda];
    }
};

template <typename T>
__global__ __launch_bounds__(CHT) void potrf_kernel(int n, Batch<T> AA, int lda, int up,
                                                    int* __restrict__ info)
{
    __shared__ T sLd[CHB][CHB + 1]; // rows of the diagonal block, one chunk of finished columns
    __shared__ T sL11[CHB][CHB + 1]; // the factored diagonal block
    __shared__ T sres[CHB]; // 1 / l_jj, or 0 after a non-positive pivot
    __shared__ int sinfo;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int64_t b = blockIdx.x;
    const LowerView<T> L{AA[b], lda, up != 0};
    if(tid == 0)
        sinfo = 0;

    for(int c0 = 0; c0 < n; c0 += CHB)
    {
        const int jb = n - c0 < CHB ? n - c0 : CHB;
        for(int r0 = c0; r0 < n; r0 += CHT) // uniform
        {
            const int i = r0 + tid; // row of this thread
            const bool act = i < n;
            T acc[CHB];
#pragma unroll
            for(int j = 0; j < CHB; ++j)
                acc[j] = (act && j < jb && c0 + j <= i) ? L(i, c0 + j) : T(0);

            // 1. products with the finished columns 0 .. c0-1, CHB at a time
            for(int k0 = 0; k0 < c0; k0 += CHB)
            {
                __syncthreads(); // previous chunk's sLd reads are done
                for(int e = tid; e < CHB * CHB; e += CHT)
                {
                    const int j = e % CHB, kk = e / CHB;
                    sLd[j][kk] = j < jb ? L(c0 + j, k0 + kk) : T(0);
                }
                __syncthreads();
                if(act)
                {
                    T li[CHB];
#pragma unroll
                    for(int kk = 0; kk < CHB; ++kk)
                        li[kk] = L(i, k0 + kk);
#pragma unroll
                    for(int j = 0; j < CHB; ++j)
                        if(c0 + j <= i) // the lower triangle only (diagonal-block rows)
#pragma unroll
                            for(int kk = 0; kk < CHB; ++kk)
                                acc[j] = fma_t(-li[kk], sLd[j][kk], acc[j]);
                }
            }

            // 2. the diagonal block: lanes 0 .. jb-1 of wavefront 0 (rows c0 + lane)
            if(r0 == c0) // uniform
            {
                if(tid < 64)
                {
#pragma unroll
                    for(int k = 0; k < CHB; ++k)
                    {
                        if(k < jb) // uniform
                        {
                            const T t = readlane_t(acc[k], k); // updated a_kk
                            T piv, r;
                            if(t <= T(0))
                            {
                                if(lane == 0 && sinfo == 0)
                                    sinfo = c0 + k + 1;
                                piv = t;
                                r = T(0);
                            }
                            else
                            {
                                piv = sqrt_t(t);
                                r = T(1) / piv;
                            }
                            if(lane == k)
                                acc[k] = piv;
                            else if(lane > k && lane < jb)
                                acc[k] *= r;
                            if(lane == 0)
                                sres[k] = r;
                            // trailing entries of the block: a_jm -= l_jk l_mk, m in (k, j]
#pragma unroll
                            for(int m = 0; m < CHB; ++m)
                                if(m > k && m < jb) // uniform
                                {
                                    const T lmk = readlane_t(acc[k], m);
                                    if(lane >= m && lane < jb)
                                        acc[m] = fma_t(-acc[k], lmk, acc[m]);
                                }
                        }
                    }
                    if(lane < jb)
                    {
#pragma unroll
                        for(int j = 0; j < CHB; ++j)
                            if(j <= lane)
                            {
                                sL11[lane][j] = acc[j];
                                L(c0 + lane, c0 + j) = acc[j];
                            }
                    }
                }
                __syncthreads();
            }

            // 3. rows below the diagonal block
            if(act && i >= c0 + jb)
            {
#pragma unroll
                for(int k = 0; k < CHB; ++k)
                {
                    if(k < jb) // uniform
                    {
                        acc[k] *= sres[k];
#pragma unroll
                        for(int m = 0; m < CHB; ++m)
                            if(m > k && m < jb)
                                acc[m] = fma_t(-acc[k], sL11[m][k], acc[m]);
                    }
                }
#pragma unroll
                for(int j = 0; j < CHB; ++j)
                    if(j < jb)
                        L(i, c0 + j) = acc[j];
            }
        }
        __syncthreads(); // the block column is written before the next one reads it
    }
    if(tid == 0)
        info[b] = sinfo;
}

// One system per workgroup (n <= NMAX): the two non-unit triangular solves
// with the factor, through the LU path's row solvers.  skip[b] != 0: B of
// system b is left unchanged (posv after a failed factorisation).
template <typename T, int NMAX, int NR>
__global__ __launch_bounds__(NMAX) void potrs_fused_kernel(int n, int nrhs, Batch<T> AA, int lda,
                                                           int up, Batch<T> BB, int ldb,
                                                           const int* __restrict__ skip)
{
    __shared__ T sx[2][GBS][NR];
    __shared__ T sL[GBS][GBS + 1];
    const int64_t b = blockIdx.x;
    if(skip && skip[b] != 0)
        return;
    const int tid = threadIdx.x;
    const MatrixBuffer<T> A(AA[b], lda, n);
    T* Bm = BB[b];
    const MatrixBuffer<T> Bbuf(Bm, ldb, nrhs);
    for(int c0 = 0; c0 < nrhs; c0 += NR)
    {
        T x[NR];
#pragma unroll
        for(int q = 0; q < NR; ++q)
        {
            const uint32_t off = (tid < n && c0 + q < nrhs)
                                     ? Bbuf.row_offset(tid) + Bbuf.col_offset(c0 + q)
                                     : MatrixBuffer<T>::OUT_OF_RANGE;
            x[q] = Bbuf.load(off, 0);
        }
        if(up)
        {
            solve_lower_rows<T, NR, true>(n, A, x, sx, sL); // U^T y = b
            solve_upper_rows<T, NR, false>(n, A, x, sx, sL); // U x = y
        }
        else
        {
            solve_lower_rows<T, NR, false, false, 1>(n, A, x, sx, sL); // L y = b
            solve_upper_rows<T, NR, true, false, 1>(n, A, x, sx, sL); // L^T x = y
        }
        if(tid < n)
        {
#pragma unroll
            for(int q = 0; q < NR; ++q)
                if(c0 + q < nrhs)
                    Bm[tid + int64_t(c0 + q) * ldb] = x[q];
        }
        __syncthreads();
    }
}

// Any order: one thread per (system, right-hand side), column-oriented
// substitution in the same order of operations.
template <typename T>
__global__ __launch_bounds__(64) void potrs_serial_kernel(int n, int nrhs, Batch<T> AA, int lda,
                                                          int up, Batch<T> BB, int ldb, int bc,
                                                          const int* __restrict__ skip)
{
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(g >= int64_t(bc) * nrhs)
        return;
    const int64_t b = g / nrhs;
    const int c = int(g % nrhs);
    if(skip && skip[b] != 0)
        return;
    const LowerView<T> L{AA[b], lda, up != 0};
    T* x = BB[b] + int64_t(c) * ldb;
    for(int k = 0; k < n; ++k)
    {
        x[k] = x[k] / L(k, k);
        for(int i = k + 1; i < n; ++i)
            x[i] = fma_t(-L(i, k), x[k], x[i]);
    }
    for(int k = n - 1; k >= 0; --k)
    {
        x[k] = x[k] / L(k, k);
        for(int i = 0; i < k; ++i)
            x[i] = fma_t(-L(k, i), x[k], x[i]);
    }
}

template <typename T>
rocblas_status potrf(hipStream_t stream, bool up, int n, Batch<T> A, int lda, int* info, int bc)
{
    if(bc <= 0)
        return rocblas_status_success;
    if(n <= 0)
        return reset_info(stream, info, bc);
    for(int64_t b0 = 0; b0 < bc; b0 += 1 << 30)
    {
        const int cnt = int((bc - b0) < (1 << 30) ? (bc - b0) : (1 << 30));
        Batch<T> Ab = A.ptrs ? Batch<T>{nullptr, A.ptrs + b0, 0}
                             : Batch<T>{A.base + b0 * A.stride, nullptr, A.stride};
        hipLaunchKernelGGL((potrf_kernel<T>), dim3(cnt), dim3(CHT), 0, stream, n, Ab, lda,
                           int(up), info + b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

template <typename T>
rocblas_status potrs(hipStream_t stream, bool up, int n, int nrhs, Batch<T> A, int lda,
                     Batch<T> B, int ldb, int bc, const int* skip)
{
    if(bc <= 0 || n <= 0 || nrhs <= 0)
        return rocblas_status_success;
    if(n <= FUSED_MAX && fits_buffer<T>(lda, n) && fits_buffer<T>(ldb, nrhs))
    {
        auto go = [&](auto nmax, auto nr) {
            constexpr int NMAX = decltype(nmax)::value;
            constexpr int NR = decltype(nr)::value;
            hipLaunchKernelGGL((potrs_fused_kernel<T, NMAX, NR>), dim3(bc), dim3(NMAX), 0, stream,
                               n, nrhs, A, lda, int(up), B, ldb, skip);
        };
        auto by_nrhs = [&](auto nmax) {
            if(nrhs <= 1)
                go(nmax, std::integral_constant<int, 1>{});
            else if(nrhs <= 4)
                go(nmax, std::integral_constant<int, 4>{});
            else
                go(nmax, std::integral_constant<int, 8>{});
        };
        if(n <= 64)
            by_nrhs(std::integral_constant<int, 64>{});
        else if(n <= 128)
            by_nrhs(std::integral_constant<int, 128>{});
        else if(n <= 256)
            by_nrhs(std::integral_constant<int, 256>{});
        else
            by_nrhs(std::integral_constant<int, 512>{});
        RSA_LAUNCH_CHECK();
        return rocblas_status_success;
    }
    const int64_t threads = int64_t(bc) * nrhs;
    hipLaunchKernelGGL((potrs_serial_kernel<T>), dim3(unsigned((threads + 63) / 64)), dim3(64), 0,
                       stream, n, nrhs, A, lda, int(up), B, ldb, bc, skip);
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

#define RSA_CHOL_INST(T)                                                                      \
    template rocblas_status potrf<T>(hipStream_t, bool, int, Batch<T>, int, int*, int);       \
    template rocblas_status potrs<T>(hipStream_t, bool, int, int, Batch<T>, int, Batch<T>, int, \
                                     int, const int*);
RSA_CHOL_INST(float)
RSA_CHOL_INST(double)

} // namespace rsa
