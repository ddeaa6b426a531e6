// adjoint.hip -- reverse-mode derivative of the batched solve X = A^-1 B.
//
// Given the LU and pivots of A (getrf/gesv output), the solution X and the
// upstream gradient dX:
//   dB = A^-T dX      -- transposed getrs reusing the stored LU and ipiv
//   dA = -dB X^T      -- batched outer product, this file
// No counterpart in the reference; the transposed solve follows getrs's
// op = T order (roclapack_getrs.hpp:134-150).
#include "internal.hpp"

namespace rsa
{

constexpr int OT = 64; // output tile
constexpr int OK = 16; // right-hand sides per LDS round

// dA[i0:i0+64, j0:j0+64] = -sum_c dB[i, c] X[j, c]; one workgroup per tile and
// matrix, each thread 16 outputs, writes coalesced along columns of dA.
template <typename T>
__global__ __launch_bounds__(256) void outer_neg_kernel(int n, int nrhs, Batch<T> DB, int lddb,
                                                        Batch<T> XX, int ldx, Batch<T> DA,
                                                        int ldda, int64_t boff)
{
    __shared__ T sB[OK][OT + 1];
    __shared__ T sX[OK][OT + 1];
    const int64_t b = blockIdx.z + boff;
    const T* dB = DB[b];
    const T* X = XX[b];
    T* dA = DA[b];
    const int i0 = blockIdx.x * OT, j0 = blockIdx.y * OT;
    const int tid = threadIdx.x;
    const int ti = tid % OT;       // row within the tile
    const int tj = (tid / OT) * 16; // first of 16 columns
    T acc[16];
#pragma unroll
    for(int q = 0; q < 16; ++q)
        acc[q] = T(0);
    for(int c0 = 0; c0 < nrhs; c0 += OK)
    {
        __syncthreads();
        for(int e = tid; e < OK * OT; e += 256)
        {
            const int r = e % OT, c = e / OT;
            const bool okc = c0 + c < nrhs;
            sB[c][r] = (okc && i0 + r < n) ? dB[(i0 + r) + int64_t(c0 + c) * lddb] : T(0);
            sX[c][r] = (okc && j0 + r < n) ? X[(j0 + r) + int64_t(c0 + c) * ldx] : T(0);
        }
        __syncthreads();
#pragma unroll
        for(int c = 0; c < OK; ++c)
        {
            const T bi = sB[c][ti];
#pragma unroll
            for(int q = 0; q < 16; ++q)
                acc[q] = fma_t(bi, sX[c][tj + q], acc[q]);
        }
    }
    const int i = i0 + ti;
    if(i < n)
    {
#pragma unroll
        for(int q = 0; q < 16; ++q)
        {
            const int j = j0 + tj + q;
            if(j < n)
                dA[i + int64_t(j) * ldda] = -acc[q];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void copy_cols_kernel(int n, int ncols, Batch<T> S, int lds,
                                                        Batch<T> D, int ldd, int64_t boff)
{
    const int64_t b = blockIdx.y + boff;
    const T* s = S[b];
    T* d = D[b];
    for(int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < int64_t(n) * ncols;
        e += int64_t(gridDim.x) * 256)
    {
        const int i = int(e % n), c = int(e / n);
        d[i + int64_t(c) * ldd] = s[i + int64_t(c) * lds];
    }
}

template <typename T>
rocblas_status gesv_adjoint(hipStream_t stream, int n, int nrhs, Batch<T> LU, int lda,
                            const int* ipiv, int64_t strideP, Batch<T> X, int ldx, Batch<T> dX,
                            int lddx, Batch<T> dB, int lddb, Batch<T> dA, int ldda, int bc)
{
    if(n == 0 || bc == 0)
        return rocblas_status_success;
    constexpr int64_t MAXYZ = 65535;
    if(nrhs > 0)
    {
        // dB = dX, then dB = A^-T dB
        for(int64_t b0 = 0; b0 < bc; b0 += MAXYZ)
        {
            const int nb = int((bc - b0) < MAXYZ ? (bc - b0) : MAXYZ);
            const int gx = ceil_div(int64_t(n) * nrhs, 256) < 64 ? ceil_div(int64_t(n) * nrhs, 256)
                                                                : 64;
            hipLaunchKernelGGL((copy_cols_kernel<T>), dim3(gx, nb), dim3(256), 0, stream, n, nrhs,
                               dX, lddx, dB, lddb, b0);
            RSA_LAUNCH_CHECK();
        }
        rocblas_status st;
        if(n <= SMALL_MAX)
            st = getrs_small<T>(stream, true, n, nrhs, LU, lda, ipiv, strideP, dB, lddb, bc, true);
        else if(n <= FUSED_MAX && fits_buffer<T>(lda, n) && fits_buffer<T>(lddb, nrhs))
            st = getrs_fused<T, T>(stream, true, n, nrhs, LU, lda, ipiv, strideP, dB, lddb, dB,
                                   lddb, bc, nullptr, false);
        else
            st = getrs_blocked<T>(stream, true, n, nrhs, LU, lda, ipiv, strideP, dB, lddb, bc, true,
                                  nullptr);
        if(st != rocblas_status_success)
            return st;
    }
    // dA = -dB X^T (zero when nrhs == 0)
    for(int64_t b0 = 0; b0 < bc; b0 += MAXYZ)
    {
        const int nb = int((bc - b0) < MAXYZ ? (bc - b0) : MAXYZ);
        dim3 grid(ceil_div(n, OT), ceil_div(n, OT), nb);
        hipLaunchKernelGGL((outer_neg_kernel<T>), grid, dim3(256), 0, stream, n, nrhs, dB, lddb, X,
                           ldx, dA, ldda, b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

template rocblas_status gesv_adjoint<float>(hipStream_t, int, int, Batch<float>, int, const int*,
                                            int64_t, Batch<float>, int, Batch<float>, int,
                                            Batch<float>, int, Batch<float>, int, int);
template rocblas_status gesv_adjoint<double>(hipStream_t, int, int, Batch<double>, int,
                                             const int*, int64_t, Batch<double>, int,
                                             Batch<double>, int, Batch<double>, int,
                                             Batch<double>, int, int);

} // namespace rsa
