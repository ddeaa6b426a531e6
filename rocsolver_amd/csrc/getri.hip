// getri.hip -- matrix inverse from the LU factorisation (getri,
// getri_outofplace and their no-pivoting variants).
//
// inv(A) is computed as the solution X of A X = I with the triangular-solve
// kernels of the LU path (the row interchanges of ipiv applied to I, then the
// unit-lower and upper solves), which is how the reference computes its
// out-of-place inverse (roclapack_getri_outofplace.hpp:100-127).  The in-place
// form (roclapack_getri.hpp:265-378 solves inv(A) L = inv(U) instead) works
// on a copy of the factors in the handle's workspace and writes X over A;
// both forms give the inverse to the reference's test tolerance.  info is the
// first zero on the diagonal of U (1-based, as the reference's trtri /
// check_singularity report it); the inverse is then undefined, as in the
// reference.
#include "internal.hpp"

namespace rsa
{

// One wavefront per matrix: the first zero diagonal entry of U.
template <typename T>
__global__ __launch_bounds__(256) void diag_zero_kernel(int n, Batch<T> A, int lda,
                                                        int* __restrict__ info, int bc)
{
    const int64_t b = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if(b >= bc)
        return;
    const int lane = threadIdx.x & 63;
    const T* a = A[b];
    int first = 0;
    for(int i0 = 0; i0 < n; i0 += 64)
    {
        const int i = i0 + lane;
        const bool z = i < n && a[i + int64_t(i) * lda] == T(0);
        const uint64_t m = __builtin_amdgcn_ballot_w64(z);
        if(m != 0) // uniform
        {
            first = i0 + int(__builtin_ctzll(m)) + 1;
            break;
        }
    }
    if(lane == 0)
        info[b] = first;
}

// C = I (n x n, leading dimension ldc) for every batch entry.
template <typename T>
__global__ __launch_bounds__(256) void set_identity_kernel(int n, Batch<T> C, int ldc, int bc)
{
    const int64_t nn = int64_t(n) * n;
    for(int64_t b = blockIdx.y; b < bc; b += gridDim.y)
    {
        T* c = C[b];
        for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nn;
            e += int64_t(gridDim.x) * blockDim.x)
        {
            const int i = int(e % n), j = int(e / n);
            c[i + int64_t(j) * ldc] = i == j ? T(1) : T(0);
        }
    }
}

// D = S (n x n) for every batch entry.
template <typename T>
__global__ __launch_bounds__(256) void copy_square_kernel(int n, Batch<T> S, int lds, Batch<T> D,
                                                          int ldd, int bc)
{
    const int64_t nn = int64_t(n) * n;
    for(int64_t b = blockIdx.y; b < bc; b += gridDim.y)
    {
        const T* s = S[b];
        T* d = D[b];
        for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nn;
            e += int64_t(gridDim.x) * blockDim.x)
        {
            const int i = int(e % n), j = int(e / n);
            d[i + int64_t(j) * ldd] = s[i + int64_t(j) * lds];
        }
    }
}

static dim3 square_grid(int n, int bc)
{
    const int64_t tiles = (int64_t(n) * n + 255) / 256;
    return dim3(unsigned(tiles < 64 ? tiles : 64), unsigned(bc < 65535 ? bc : 65535));
}

template <typename T>
rocblas_status getri_solve(hipStream_t stream, int n, Batch<T> LU, int ldlu, const int* ipiv,
                           int64_t strideP, Batch<T> X, int ldx, int* info, int bc, bool pivot)
{
    if(n <= 0 || bc <= 0)
        return rocblas_status_success;
    hipLaunchKernelGGL((diag_zero_kernel<T>), dim3((bc + 3) / 4), dim3(256), 0, stream, n, LU,
                       ldlu, info, bc);
    RSA_LAUNCH_CHECK();
    hipLaunchKernelGGL((set_identity_kernel<T>), square_grid(n, bc), dim3(256), 0, stream, n, X,
                       ldx, bc);
    RSA_LAUNCH_CHECK();
    if(n <= SMALL_MAX)
        return getrs_small<T>(stream, false, n, n, LU, ldlu, ipiv, strideP, X, ldx, bc, pivot);
    if(pivot && n <= FUSED_MAX && fits_buffer<T>(ldlu, n) && fits_buffer<T>(ldx, n))
        return getrs_fused<T, T>(stream, false, n, n, LU, ldlu, ipiv, strideP, X, ldx, X, ldx, bc,
                                 nullptr, false);
    return getrs_blocked<T>(stream, false, n, n, LU, ldlu, ipiv, strideP, X, ldx, bc, pivot,
                            nullptr);
}

template <typename T>
rocblas_status copy_square(hipStream_t stream, int n, Batch<T> S, int lds, Batch<T> D, int ldd,
                           int bc)
{
    if(n <= 0 || bc <= 0)
        return rocblas_status_success;
    hipLaunchKernelGGL((copy_square_kernel<T>), square_grid(n, bc), dim3(256), 0, stream, n, S,
                       lds, D, ldd, bc);
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

#define RSA_GETRI_INST(T)                                                                     \
    template rocblas_status getri_solve<T>(hipStream_t, int, Batch<T>, int, const int*,       \
                                           int64_t, Batch<T>, int, int*, int, bool);          \
    template rocblas_status copy_square<T>(hipStream_t, int, Batch<T>, int, Batch<T>, int, int);
RSA_GETRI_INST(float)
RSA_GETRI_INST(double)

} // namespace rsa
