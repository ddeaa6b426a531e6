// api.cpp -- extern "C" entry points (the drop-in boundary).
//
// Each *_impl reproduces the reference impl's control flow:
//   null handle -> invalid_handle;  argCheck in the reference order
//   (values, then sizes, then -- unless in device-memory size-query mode --
//   pointers);  size query -> report the workspace (the fused LU keeps its
//   inverted diagonal blocks there; the other paths need none);
//   quick returns;  then the work, queued on the handle's stream.
// References: roclapack_getrf_strided_batched.cpp:7-84 (getrf impl),
// roclapack_getf2.hpp:488-517 (getf2/getrf argCheck), roclapack_getrs.hpp:17-49
// (getrs argCheck) and :99-107 (quick return), roclapack_gesv.hpp:17-46
// (gesv argCheck) and :140-152 (info reset and quick return),
// rocauxiliary_laswp.hpp:66-94 and rocauxiliary_laswp.cpp:7-42 (laswp).
#include "../../include/rocsolver_amd.h"
#include "internal.hpp"
#include "logging.hpp"

#include <rocblas/internal/rocblas_device_malloc.hpp>

#include <algorithm>
#include <cstring>

using namespace rsa;

namespace
{

// the work needs no device workspace: a size query records zero bytes
rocblas_status report_no_workspace(rocblas_handle handle)
{
    return rocblas_set_optimal_device_memory_size(handle, size_t(0));
}

template <typename T>
rocblas_status getrf_impl(rocblas_handle handle, int m, int n, Batch<T> A, bool hasA, int lda,
                          int* ipiv, int64_t strideP, int* info, int bc, bool pivot)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(m < 0 || n < 0 || lda < m || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    const int64_t mn = int64_t(m) * n;
    if(!query && ((mn && !hasA) || (mn && pivot && !ipiv) || (bc && !info)))
        return rocblas_status_invalid_pointer;
    // the fused path keeps the inverted diagonal blocks of L in workspace
    const bool fused = pivot && m == n && n > SMALL_MAX && n <= FUSED_MAX
                       && fits_buffer<T>(lda, n);
    const size_t ws = fused && bc > 0 ? lu_fused_workspace_bytes<T>(n, bc) : 0;
    if(query)
        return ws ? rocblas_set_optimal_device_memory_size(handle, ws)
                  : report_no_workspace(handle);

    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(bc == 0)
        return rocblas_status_success;
    if(m == 0 || n == 0)
        return reset_info(stream, info, bc);
    if(m <= SMALL_MAX && n <= SMALL_MAX)
    {
        RSA_LOG_STAGE(handle, "getrf_small", "m", m, "n", n, "pivot", int(pivot), "bc", bc);
        return getrf_small<T>(stream, m, n, A, lda, ipiv, strideP, info, bc, pivot);
    }
    if(fused)
    {
        RSA_LOG_STAGE(handle, "getrf_fused", "n", n, "bc", bc);
        rocblas_device_malloc mem(handle, ws);
        if(!mem)
            return rocblas_status_memory_error;
        return lu_fused<T, T>(stream, n, A, lda, A, lda, ipiv, strideP, info, bc,
                              strided(static_cast<T*>(static_cast<void*>(mem)),
                                      lu_fused_inv_elems(n)));
    }
    RSA_LOG_STAGE(handle, "getrf_blocked", "m", m, "n", n, "pivot", int(pivot), "bc", bc);
    return getrf_blocked<T>(stream, m, n, A, lda, ipiv, strideP, info, bc, pivot);
}

rocblas_status check_trans(rocblas_operation trans, bool* transposed)
{
    if(trans != rocblas_operation_none && trans != rocblas_operation_transpose
       && trans != rocblas_operation_conjugate_transpose)
        return rocblas_status_invalid_value;
    *transposed = trans != rocblas_operation_none; // real data: C == T
    return rocblas_status_continue;
}

template <typename T>
rocblas_status getrs_impl(rocblas_handle handle, rocblas_operation trans, int n, int nrhs,
                          Batch<T> A, bool hasA, int lda, const int* ipiv, int64_t strideP,
                          Batch<T> B, bool hasB, int ldb, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    bool tr;
    rocblas_status st = check_trans(trans, &tr);
    if(st != rocblas_status_continue)
        return st;
    if(n < 0 || nrhs < 0 || lda < n || ldb < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && ((n && !hasA) || (n && !ipiv) || (int64_t(nrhs) * n && !hasB)))
        return rocblas_status_invalid_pointer;
    if(query)
        return report_no_workspace(handle);

    if(n == 0 || nrhs == 0 || bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(n <= SMALL_MAX)
    {
        RSA_LOG_STAGE(handle, "getrs_small", "n", n, "nrhs", nrhs, "trans", trans, "bc", bc);
        return getrs_small<T>(stream, tr, n, nrhs, A, lda, ipiv, strideP, B, ldb, bc, true);
    }
    if(n <= FUSED_MAX && fits_buffer<T>(lda, n) && fits_buffer<T>(ldb, nrhs))
    {
        RSA_LOG_STAGE(handle, "getrs_fused", "n", n, "nrhs", nrhs, "trans", trans, "bc", bc);
        return getrs_fused<T, T>(stream, tr, n, nrhs, A, lda, ipiv, strideP, B, ldb, B, ldb, bc,
                                 nullptr, false);
    }
    RSA_LOG_STAGE(handle, "getrs_blocked", "n", n, "nrhs", nrhs, "trans", trans, "bc", bc);
    return getrs_blocked<T>(stream, tr, n, nrhs, A, lda, ipiv, strideP, B, ldb, bc, true,
                            nullptr);
}

template <typename T>
rocblas_status gesv_impl(rocblas_handle handle, int n, int nrhs, Batch<T> A, bool hasA, int lda,
                         int* ipiv, int64_t strideP, Batch<T> B, bool hasB, int ldb, int* info,
                         int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n < 0 || nrhs < 0 || lda < n || ldb < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query
       && ((n && !hasA) || (n && !ipiv) || (int64_t(nrhs) * n && !hasB) || (bc && !info)))
        return rocblas_status_invalid_pointer;
    const bool fused = n > SMALL_MAX && n <= FUSED_MAX && fits_buffer<T>(lda, n)
                       && fits_buffer<T>(ldb, nrhs);
    const size_t ws = fused && nrhs > 0 && bc > 0 ? lu_fused_workspace_bytes<T>(n, bc) : 0;
    if(query)
        return ws ? rocblas_set_optimal_device_memory_size(handle, ws)
                  : report_no_workspace(handle);

    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(n == 0 || nrhs == 0)
        return reset_info(stream, info, bc);
    if(n <= SMALL_MAX)
        return gesv_small<T>(stream, n, nrhs, A, lda, ipiv, strideP, B, ldb, info, bc);
    if(fused)
    {
        rocblas_device_malloc mem(handle, ws);
        if(!mem)
            return rocblas_status_memory_error;
        rocblas_status st
            = lu_fused<T, T>(stream, n, A, lda, A, lda, ipiv, strideP, info, bc,
                             strided(static_cast<T*>(static_cast<void*>(mem)),
                                     lu_fused_inv_elems(n)));
        if(st != rocblas_status_success)
            return st;
        return getrs_fused<T, T>(stream, false, n, nrhs, A, lda, ipiv, strideP, B, ldb, B, ldb,
                                 bc, info, false);
    }
    rocblas_status st = getrf_blocked<T>(stream, n, n, A, lda, ipiv, strideP, info, bc, true);
    if(st != rocblas_status_success)
        return st;
    return getrs_blocked<T>(stream, false, n, nrhs, A, lda, ipiv, strideP, B, ldb, bc, true,
                            info);
}

template <typename T>
rocblas_status laswp_impl(rocblas_handle handle, int n, T* A, int lda, int k1, int k2,
                          const int* ipiv, int incx)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n < 0 || lda < 1 || !incx || k1 < 1 || k2 < 1 || k2 < k1)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && ((n && !A) || !ipiv))
        return rocblas_status_invalid_pointer;
    if(query)
        return rocblas_status_size_unchanged;
    if(n == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    // pivots are read as ipiv[k1 + (i-k1)*|incx| - 1], as the reference does
    return laswp<T>(stream, n, strided(A, 0), lda, k1, k2, ipiv, 0, incx, 1, nullptr);
}

// getri / getri_outofplace (and the no-pivoting forms): argument checks of
// roclapack_getri.hpp:216-242 and roclapack_getri_outofplace.hpp:40-67
// (sizes, then pointers unless querying), quick returns of :280-295.  The
// in-place form takes n*n*bc elements of workspace for a copy of the factors.
template <typename T>
rocblas_status getri_impl(rocblas_handle handle, int n, Batch<T> A, bool hasA, int lda,
                          int* ipiv, int64_t strideP, bool outofplace, Batch<T> C, bool hasC,
                          int ldc, int* info, int bc, bool pivot)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n < 0 || lda < n || (outofplace && ldc < n) || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query
       && ((n && !hasA) || (outofplace && n && !hasC) || (n && pivot && !ipiv) || (bc && !info)))
        return rocblas_status_invalid_pointer;
    const size_t ws = (!outofplace && n > 0 && bc > 0) ? sizeof(T) * size_t(n) * n * size_t(bc) : 0;
    if(query)
        return ws ? rocblas_set_optimal_device_memory_size(handle, ws) : report_no_workspace(handle);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(n == 0)
        return reset_info(stream, info, bc);
    if(outofplace)
        return getri_solve<T>(stream, n, A, lda, ipiv, strideP, C, ldc, info, bc, pivot);
    rocblas_device_malloc mem(handle, ws);
    if(!mem)
        return rocblas_status_memory_error;
    Batch<T> W = strided(static_cast<T*>(static_cast<void*>(mem)), int64_t(n) * n);
    rocblas_status st = copy_square<T>(stream, n, A, lda, W, n, bc);
    if(st != rocblas_status_success)
        return st;
    return getri_solve<T>(stream, n, W, n, ipiv, strideP, A, lda, info, bc, pivot);
}

// Cholesky family: argument checks of roclapack_potf2.hpp:95-121 (potf2 /
// potrf), roclapack_potrs.hpp:20-46 and roclapack_posv.hpp:20-48: uplo value,
// then sizes, then (unless querying) pointers; no workspace is needed.
inline rocblas_status check_uplo(rocblas_fill uplo, bool* up)
{
    if(uplo != rocblas_fill_upper && uplo != rocblas_fill_lower)
        return rocblas_status_invalid_value;
    *up = uplo == rocblas_fill_upper;
    return rocblas_status_continue;
}

template <typename T>
rocblas_status potrf_impl(rocblas_handle handle, rocblas_fill uplo, int n, Batch<T> A, bool hasA,
                          int lda, int* info, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    bool up;
    rocblas_status st = check_uplo(uplo, &up);
    if(st != rocblas_status_continue)
        return st;
    if(n < 0 || lda < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && ((n && !hasA) || (bc && !info)))
        return rocblas_status_invalid_pointer;
    if(query)
        return report_no_workspace(handle);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    return potrf<T>(stream, up, n, A, lda, info, bc);
}

template <typename T>
rocblas_status potrs_impl(rocblas_handle handle, rocblas_fill uplo, int n, int nrhs, Batch<T> A,
                          bool hasA, int lda, Batch<T> B, bool hasB, int ldb, int* info, int bc,
                          bool posv)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    bool up;
    rocblas_status st = check_uplo(uplo, &up);
    if(st != rocblas_status_continue)
        return st;
    if(n < 0 || nrhs < 0 || lda < n || ldb < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && ((n && !hasA) || (int64_t(nrhs) * n && !hasB) || (posv && bc && !info)))
        return rocblas_status_invalid_pointer;
    if(query)
        return report_no_workspace(handle);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(!posv)
        return potrs<T>(stream, up, n, nrhs, A, lda, B, ldb, bc, nullptr);
    // posv: info reset, then nothing more when A or B is empty
    // (roclapack_posv.hpp:135-142); B is left unchanged where info != 0
    if(n == 0 || nrhs == 0)
        return reset_info(stream, info, bc);
    st = potrf<T>(stream, up, n, A, lda, info, bc);
    if(st != rocblas_status_success)
        return st;
    return potrs<T>(stream, up, n, nrhs, A, lda, B, ldb, bc, info);
}

constexpr int DSGESV_ITERMAX = 30; // LAPACK dsgesv ITERMAX

// D[b][i, c] = S[b][i, c] for i < m, c < ncols; batch entries boff + blockIdx.y
__global__ void copy_block_kernel(int m, int ncols, Batch<double> S, int lds, Batch<double> D,
                                  int ldd, int64_t boff)
{
    const int64_t b = blockIdx.y + boff;
    const double* s = S[b];
    double* d = D[b];
    for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < int64_t(m) * ncols;
        e += int64_t(gridDim.x) * blockDim.x)
    {
        const int i = int(e % m), c = int(e / m);
        d[i + int64_t(c) * ldd] = s[i + int64_t(c) * lds];
    }
}

__global__ void mark_fp64_only_kernel(int* iter, double* berr, int bc)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if(b < bc)
    {
        iter[b] = -1;
        if(berr)
            berr[b] = -1.0;
    }
}

static rocblas_status copy_block(hipStream_t stream, int m, int ncols, Batch<double> S, int lds,
                                 Batch<double> D, int ldd, int bc)
{
    const int gx = int(std::min<int64_t>((int64_t(m) * ncols + 255) / 256, 1024));
    for(int64_t b0 = 0; b0 < bc; b0 += 65535)
    {
        const int nb = int(std::min<int64_t>(bc - b0, 65535));
        hipLaunchKernelGGL(copy_block_kernel, dim3(gx, nb), dim3(256), 0, stream, m, ncols, S,
                           lds, D, ldd, b0);
        RSA_LAUNCH_CHECK();
    }
    return rocblas_status_success;
}

// Orders beyond the fused kernels (or a matrix beyond their 32-bit buffer
// offsets): every system is solved by the FP64 LU of the blocked path on a
// copy of A in the workspace, and iter = -1, LAPACK DSGESV's code for "fell
// back to full precision for implementation- or machine-specific reasons";
// berr = -1 as on every FP64 path.  X = B where info != 0.
static rocblas_status dsgesv_fp64_only(hipStream_t stream, int n, int nrhs, Batch<double> A,
                                       int lda, Batch<double> B, int ldb, Batch<double> X,
                                       int ldx, int* iter, int* info, double* berr, int bc,
                                       void* work)
{
    const int64_t slot = int64_t(n) * n;
    double* lu = static_cast<double*>(work);
    int* ipiv = reinterpret_cast<int*>(
        static_cast<char*>(work) + (sizeof(double) * slot * bc + 255) / 256 * 256);
    const Batch<double> LU = strided(lu, slot);
    rocblas_status st = copy_block(stream, n, n, A, lda, LU, n, bc);
    if(st == rocblas_status_success)
        st = copy_block(stream, n, nrhs, B, ldb, X, ldx, bc);
    if(st == rocblas_status_success)
        st = getrf_blocked<double>(stream, n, n, LU, n, ipiv, n, info, bc, true);
    if(st == rocblas_status_success)
        st = getrs_blocked<double>(stream, false, n, nrhs, LU, n, ipiv, n, X, ldx, bc, true,
                                   info);
    if(st != rocblas_status_success)
        return st;
    hipLaunchKernelGGL(mark_fp64_only_kernel, dim3((bc + 255) / 256), dim3(256), 0, stream, iter,
                       berr, bc);
    RSA_LAUNCH_CHECK();
    return rocblas_status_success;
}

rocblas_status dsgesv_impl(rocblas_handle handle, int factor, int n, int nrhs, const double* A,
                           int lda, int64_t strideA, const double* B, int ldb, int64_t strideB,
                           double* X, int ldx, int64_t strideX, int* iter, int* info,
                           double* berr, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(factor < 0 || factor > 2)
        return rocblas_status_invalid_value;
    if(n < 0 || nrhs < 0 || lda < n || ldb < n || ldx < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    const int64_t nn = int64_t(n) * nrhs;
    if(!query && ((n && !A) || (nn && (!B || !X)) || (bc && (!iter || !info))))
        return rocblas_status_invalid_pointer;
    const bool work_needed = n > 0 && nrhs > 0 && bc > 0;
    const size_t ws = work_needed ? dsgesv_workspace_bytes(n, nrhs, bc, DSGESV_ITERMAX) : 0;
    if(query)
        return rocblas_set_optimal_device_memory_size(handle, ws);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(!work_needed)
    {
        rocblas_status st = reset_info(stream, info, bc);
        return st != rocblas_status_success ? st : reset_info(stream, iter, bc);
    }
    rocblas_device_malloc mem(handle, ws);
    if(!mem)
        return rocblas_status_memory_error;
    if(n > FUSED_MAX || !fits_buffer<double>(lda, n)
       || !fits_buffer<double>(ldb, nrhs) || !fits_buffer<double>(ldx, nrhs))
        return dsgesv_fp64_only(stream, n, nrhs, strided(const_cast<double*>(A), strideA), lda,
                                strided(const_cast<double*>(B), strideB), ldb,
                                strided(X, strideX), ldx, iter, info, berr, bc,
                                static_cast<void*>(mem));
    return dsgesv_refine(stream, factor, n, nrhs, strided(const_cast<double*>(A), strideA), lda,
                         strided(const_cast<double*>(B), strideB), ldb, strided(X, strideX), ldx,
                         iter, info, berr, bc, DSGESV_ITERMAX, static_cast<void*>(mem));
}

__global__ void convert_d2s_kernel(int m, int n, Batch<double> A, int lda, Batch<float> O, int ldo)
{
    const int64_t b = blockIdx.y;
    const double* a = A[b];
    float* o = O[b];
    for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < int64_t(m) * n;
        e += int64_t(gridDim.x) * blockDim.x)
    {
        const int i = int(e % m), j = int(e / m);
        o[i + int64_t(j) * ldo] = float(a[i + int64_t(j) * lda]);
    }
}

rocblas_status dsgetrf_impl(rocblas_handle handle, int n, const double* A, int lda,
                            int64_t strideA, float* LU, int ldlu, int64_t strideLU, int* ipiv,
                            int64_t strideP, int* info, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n < 0 || lda < n || ldlu < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    const int64_t nn = int64_t(n) * n;
    if(!query && ((nn && (!A || !LU || !ipiv)) || (bc && !info)))
        return rocblas_status_invalid_pointer;
    const bool fused = n > SMALL_MAX && n <= FUSED_MAX && fits_buffer<double>(lda, n)
                       && fits_buffer<float>(ldlu, n);
    const size_t ws = fused && bc > 0 ? lu_fused_workspace_bytes<float>(n, bc) : 0;
    if(query)
        return ws ? rocblas_set_optimal_device_memory_size(handle, ws)
                  : report_no_workspace(handle);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    if(n == 0)
        return reset_info(stream, info, bc);
    Batch<double> Ab = strided(const_cast<double*>(A), strideA);
    Batch<float> Lb = strided(LU, strideLU);
    if(fused)
    {
        rocblas_device_malloc mem(handle, ws);
        if(!mem)
            return rocblas_status_memory_error;
        return lu_fused<double, float>(stream, n, Ab, lda, Lb, ldlu, ipiv, strideP, info, bc,
                                       strided(static_cast<float*>(static_cast<void*>(mem)),
                                               lu_fused_inv_elems(n)));
    }
    // other orders: round to FP32 into LU, then the FP32 factorisation in place
    for(int64_t b0 = 0; b0 < bc; b0 += 65535)
    {
        const int nb = int((bc - b0) < 65535 ? (bc - b0) : 65535);
        Batch<double> Ai = strided(const_cast<double*>(A) + b0 * strideA, strideA);
        Batch<float> Li = strided(LU + b0 * strideLU, strideLU);
        hipLaunchKernelGGL(convert_d2s_kernel, dim3(ceil_div(nn, 256) < 64 ? ceil_div(nn, 256) : 64, nb),
                           dim3(256), 0, stream, n, n, Ai, lda, Li, ldlu);
        RSA_LAUNCH_CHECK();
    }
    return getrf_impl<float>(handle, n, n, Lb, true, ldlu, ipiv, strideP, info, bc, true);
}

template <typename T>
rocblas_status adjoint_impl(rocblas_handle handle, int n, int nrhs, const T* LU, int lda,
                            int64_t strideA, const int* ipiv, int64_t strideP, const T* X,
                            int ldx, int64_t strideX, const T* dX, int lddx, int64_t strideDX,
                            T* dB, int lddb, int64_t strideDB, T* dA, int ldda, int64_t strideDA,
                            int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n < 0 || nrhs < 0 || lda < n || ldx < n || lddx < n || lddb < n || ldda < n || bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    const int64_t nn = int64_t(n) * nrhs;
    if(!query && ((n && (!LU || !ipiv || !dA)) || (nn && (!X || !dX || !dB))))
        return rocblas_status_invalid_pointer;
    if(query)
        return report_no_workspace(handle);
    if(n == 0 || bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    auto cb = [](const T* p, int64_t s) { return strided(const_cast<T*>(p), s); };
    return gesv_adjoint<T>(stream, n, nrhs, cb(LU, strideA), lda, ipiv, strideP, cb(X, strideX),
                           ldx, cb(dX, strideDX), lddx, strided(dB, strideDB), lddb,
                           strided(dA, strideDA), ldda, bc);
}

template <typename T>
rocblas_status vbatched_impl(rocblas_handle handle, int op, rocblas_operation trans,
                             const int* n, const int* nrhs, T* const* A, const int* lda,
                             int* const* ipiv, T* const* B, const int* ldb, int* info, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    bool tr = false;
    if(op == 1)
    {
        rocblas_status st = check_trans(trans, &tr);
        if(st != rocblas_status_continue)
            return st;
    }
    if(bc < 0)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && bc > 0
       && (!n || !lda || !A || !ipiv || (op != 0 && (!nrhs || !B || !ldb))
           || (op != 1 && !info)))
        return rocblas_status_invalid_pointer;
    const size_t ws = bc > 0 ? vbatched_workspace_bytes(bc) : 0;
    if(query)
        return rocblas_set_optimal_device_memory_size(handle, ws);
    if(bc == 0)
        return rocblas_status_success;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    rocblas_device_malloc mem(handle, ws);
    if(!mem)
        return rocblas_status_memory_error;
    return vbatched<T>(stream, op, tr, n, nrhs, A, lda, ipiv, B, ldb, info, bc,
                       static_cast<void*>(mem));
}

// gesv of a variable-size FP64 batch with the work queue in the library's
// static device array instead of handle workspace: no allocation at all.
rocblas_status dgesv_vbatched_noalloc_impl(rocblas_handle handle, const int* n, const int* nrhs,
                                           double* const* A, const int* lda, int* const* ipiv,
                                           double* const* B, const int* ldb, int* info, int bc)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(bc < 0 || bc > VB_NOALLOC_MAX_BATCH)
        return rocblas_status_invalid_size;
    const bool query = rocblas_is_device_memory_size_query(handle);
    if(!query && bc > 0 && (!n || !lda || !A || !ipiv || !nrhs || !B || !ldb || !info))
        return rocblas_status_invalid_pointer;
    if(query)
        return report_no_workspace(handle);
    if(bc == 0)
        return rocblas_status_success;
    void* work = vbatched_noalloc_scratch();
    if(!work)
        return rocblas_status_internal_error;
    hipStream_t stream;
    rocblas_get_stream(handle, &stream);
    return vbatched<double>(stream, 2, false, n, nrhs, A, lda, ipiv, B, ldb, info, bc, work);
}

} // namespace

#define RSA_GETRF_FAMILY(P, T)                                                                \
    rocblas_status rocsolver_##P##getrf(rocblas_handle handle, const rocblas_int m,             \
                                        const rocblas_int n, T* A, const rocblas_int lda,       \
                                        rocblas_int* ipiv, rocblas_int* info)                   \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf", "-m", m, "-n", n, "--lda", lda);                    \
        return getrf_impl<T>(handle, m, n, strided(A, 0), A != nullptr, lda, ipiv, 0, info, 1, \
                             true);                                                             \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrf_batched(rocblas_handle handle, const rocblas_int m,     \
                                                const rocblas_int n, T* const A[],              \
                                                const rocblas_int lda, rocblas_int* ipiv,       \
                                                const rocblas_stride strideP,                   \
                                                rocblas_int* info, const rocblas_int bc)        \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf_batched", "-m", m, "-n", n, "--lda", lda, "--strideP", strideP, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, arrayed<T>(A), A != nullptr, lda, ipiv, strideP,     \
                             info, bc, true);                                                   \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrf_strided_batched(                                        \
        rocblas_handle handle, const rocblas_int m, const rocblas_int n, T* A,                  \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* ipiv,                 \
        const rocblas_stride strideP, rocblas_int* info, const rocblas_int bc)                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf_strided_batched", "-m", m, "-n", n, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, strided(A, strideA), A != nullptr, lda, ipiv,        \
                             strideP, info, bc, true);                                          \
    }                                                                                           \
    rocblas_status rocsolver_##P##getf2(rocblas_handle handle, const rocblas_int m,             \
                                        const rocblas_int n, T* A, const rocblas_int lda,       \
                                        rocblas_int* ipiv, rocblas_int* info)                   \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getf2", "-m", m, "-n", n, "--lda", lda);                    \
        return getrf_impl<T>(handle, m, n, strided(A, 0), A != nullptr, lda, ipiv, 0, info, 1, \
                             true);                                                             \
    }                                                                                           \
    rocblas_status rocsolver_##P##getf2_batched(rocblas_handle handle, const rocblas_int m,     \
                                                const rocblas_int n, T* const A[],              \
                                                const rocblas_int lda, rocblas_int* ipiv,       \
                                                const rocblas_stride strideP,                   \
                                                rocblas_int* info, const rocblas_int bc)        \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getf2_batched", "-m", m, "-n", n, "--lda", lda, "--strideP", strideP, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, arrayed<T>(A), A != nullptr, lda, ipiv, strideP,     \
                             info, bc, true);                                                   \
    }                                                                                           \
    rocblas_status rocsolver_##P##getf2_strided_batched(                                        \
        rocblas_handle handle, const rocblas_int m, const rocblas_int n, T* A,                  \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* ipiv,                 \
        const rocblas_stride strideP, rocblas_int* info, const rocblas_int bc)                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getf2_strided_batched", "-m", m, "-n", n, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, strided(A, strideA), A != nullptr, lda, ipiv,        \
                             strideP, info, bc, true);                                          \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrf_npvt(rocblas_handle handle, const rocblas_int m,        \
                                             const rocblas_int n, T* A, const rocblas_int lda,  \
                                             rocblas_int* info)                                 \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf_npvt", "-m", m, "-n", n, "--lda", lda);               \
        return getrf_impl<T>(handle, m, n, strided(A, 0), A != nullptr, lda, nullptr, 0, info,  \
                             1, false);                                                         \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrf_npvt_batched(rocblas_handle handle, const rocblas_int m, \
                                                     const rocblas_int n, T* const A[],         \
                                                     const rocblas_int lda, rocblas_int* info,  \
                                                     const rocblas_int bc)                      \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf_npvt_batched", "-m", m, "-n", n, "--lda", lda, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, arrayed<T>(A), A != nullptr, lda, nullptr, 0, info,  \
                             bc, false);                                                        \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrf_npvt_strided_batched(                                   \
        rocblas_handle handle, const rocblas_int m, const rocblas_int n, T* A,                  \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* info,                 \
        const rocblas_int bc)                                                                   \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrf_npvt_strided_batched", "-m", m, "-n", n, "--lda", lda, "--strideA", strideA, "--batch_count", bc);\
        return getrf_impl<T>(handle, m, n, strided(A, strideA), A != nullptr, lda, nullptr, 0,  \
                             info, bc, false);                                                  \
    }

#define RSA_GETRS_GESV_FAMILY(P, T)                                                             \
    rocblas_status rocsolver_##P##getrs(rocblas_handle handle, const rocblas_operation trans,   \
                                        const rocblas_int n, const rocblas_int nrhs, T* A,      \
                                        const rocblas_int lda, const rocblas_int* ipiv, T* B,   \
                                        const rocblas_int ldb)                                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrs", "--trans", trans, "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb);\
        return getrs_impl<T>(handle, trans, n, nrhs, strided(A, 0), A != nullptr, lda, ipiv, 0, \
                             strided(B, 0), B != nullptr, ldb, 1);                              \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrs_batched(                                                \
        rocblas_handle handle, const rocblas_operation trans, const rocblas_int n,              \
        const rocblas_int nrhs, T* const A[], const rocblas_int lda, const rocblas_int* ipiv,    \
        const rocblas_stride strideP, T* const B[], const rocblas_int ldb, const rocblas_int bc) \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrs_batched", "--trans", trans, "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideP", strideP, "--ldb", ldb, "--batch_count", bc);\
        return getrs_impl<T>(handle, trans, n, nrhs, arrayed<T>(A), A != nullptr, lda, ipiv,    \
                             strideP, arrayed<T>(B), B != nullptr, ldb, bc);                    \
    }                                                                                           \
    rocblas_status rocsolver_##P##getrs_strided_batched(                                        \
        rocblas_handle handle, const rocblas_operation trans, const rocblas_int n,              \
        const rocblas_int nrhs, T* A, const rocblas_int lda, const rocblas_stride strideA,      \
        const rocblas_int* ipiv, const rocblas_stride strideP, T* B, const rocblas_int ldb,     \
        const rocblas_stride strideB, const rocblas_int bc)                                     \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getrs_strided_batched", "--trans", trans, "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--ldb", ldb, "--strideB", strideB, "--batch_count", bc);\
        return getrs_impl<T>(handle, trans, n, nrhs, strided(A, strideA), A != nullptr, lda,    \
                             ipiv, strideP, strided(B, strideB), B != nullptr, ldb, bc);        \
    }                                                                                           \
    rocblas_status rocsolver_##P##gesv(rocblas_handle handle, const rocblas_int n,              \
                                       const rocblas_int nrhs, T* A, const rocblas_int lda,     \
                                       rocblas_int* ipiv, T* B, const rocblas_int ldb,          \
                                       rocblas_int* info)                                       \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "gesv", "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb);\
        return gesv_impl<T>(handle, n, nrhs, strided(A, 0), A != nullptr, lda, ipiv, 0,         \
                            strided(B, 0), B != nullptr, ldb, info, 1);                         \
    }                                                                                           \
    rocblas_status rocsolver_##P##gesv_batched(                                                 \
        rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, T* const A[],       \
        const rocblas_int lda, rocblas_int* ipiv, const rocblas_stride strideP, T* const B[],   \
        const rocblas_int ldb, rocblas_int* info, const rocblas_int bc)                         \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "gesv_batched", "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideP", strideP, "--ldb", ldb, "--batch_count", bc);\
        return gesv_impl<T>(handle, n, nrhs, arrayed<T>(A), A != nullptr, lda, ipiv, strideP,   \
                            arrayed<T>(B), B != nullptr, ldb, info, bc);                        \
    }                                                                                           \
    rocblas_status rocsolver_##P##gesv_strided_batched(                                         \
        rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, T* A,               \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* ipiv,                 \
        const rocblas_stride strideP, T* B, const rocblas_int ldb,                              \
        const rocblas_stride strideB, rocblas_int* info, const rocblas_int bc)                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "gesv_strided_batched", "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--ldb", ldb, "--strideB", strideB, "--batch_count", bc);\
        return gesv_impl<T>(handle, n, nrhs, strided(A, strideA), A != nullptr, lda, ipiv,      \
                            strideP, strided(B, strideB), B != nullptr, ldb, info, bc);         \
    }                                                                                           \
    rocblas_status rocsolver_##P##laswp(rocblas_handle handle, const rocblas_int n, T* A,       \
                                        const rocblas_int lda, const rocblas_int k1,            \
                                        const rocblas_int k2, const rocblas_int* ipiv,          \
                                        const rocblas_int incx)                                 \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "laswp", "-n", n, "--lda", lda, "--k1", k1, "--k2", k2, "--incx", incx);\
        return laswp_impl<T>(handle, n, A, lda, k1, k2, ipiv, incx);                            \
    }

extern "C" {

RSA_GETRF_FAMILY(s, float)
RSA_GETRF_FAMILY(d, double)
RSA_GETRS_GESV_FAMILY(s, float)
RSA_GETRS_GESV_FAMILY(d, double)

rocblas_status rocsolver_amd_dsgesv_strided_batched(
    rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, const double* A,
    const rocblas_int lda, const rocblas_stride strideA, const double* B, const rocblas_int ldb,
    const rocblas_stride strideB, double* X, const rocblas_int ldx, const rocblas_stride strideX,
    rocblas_int* iter, rocblas_int* info, double* berr, const rocblas_int batch_count)
{
    return dsgesv_impl(handle, 0, n, nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX, iter,
                       info, berr, batch_count);
}

rocblas_status rocsolver_amd_dxgesv_strided_batched(
    rocblas_handle handle, const rocblas_int factor, const rocblas_int n, const rocblas_int nrhs,
    const double* A, const rocblas_int lda, const rocblas_stride strideA, const double* B,
    const rocblas_int ldb, const rocblas_stride strideB, double* X, const rocblas_int ldx,
    const rocblas_stride strideX, rocblas_int* iter, rocblas_int* info, double* berr,
    const rocblas_int batch_count)
{
    return dsgesv_impl(handle, factor, n, nrhs, A, lda, strideA, B, ldb, strideB, X, ldx, strideX,
                       iter, info, berr, batch_count);
}

rocblas_status rocsolver_amd_dsgetrf_strided_batched(
    rocblas_handle handle, const rocblas_int n, const double* A, const rocblas_int lda,
    const rocblas_stride strideA, float* LU, const rocblas_int ldlu, const rocblas_stride strideLU,
    rocblas_int* ipiv, const rocblas_stride strideP, rocblas_int* info,
    const rocblas_int batch_count)
{
    return dsgetrf_impl(handle, n, A, lda, strideA, LU, ldlu, strideLU, ipiv, strideP, info,
                        batch_count);
}

#define RSA_ADJOINT(P, T)                                                                       \
    rocblas_status rocsolver_amd_##P##gesv_adjoint_strided_batched(                             \
        rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, const T* LU,        \
        const rocblas_int lda, const rocblas_stride strideA, const rocblas_int* ipiv,           \
        const rocblas_stride strideP, const T* X, const rocblas_int ldx,                        \
        const rocblas_stride strideX, const T* dX, const rocblas_int lddx,                      \
        const rocblas_stride strideDX, T* dB, const rocblas_int lddb,                           \
        const rocblas_stride strideDB, T* dA, const rocblas_int ldda,                           \
        const rocblas_stride strideDA, const rocblas_int batch_count)                           \
    {                                                                                           \
        return adjoint_impl<T>(handle, n, nrhs, LU, lda, strideA, ipiv, strideP, X, ldx,        \
                               strideX, dX, lddx, strideDX, dB, lddb, strideDB, dA, ldda,       \
                               strideDA, batch_count);                                          \
    }
RSA_ADJOINT(s, float)
RSA_ADJOINT(d, double)

#define RSA_GETRI(P, T)                                                                         \
    rocblas_status rocsolver_##P##getri(rocblas_handle handle, const rocblas_int n, T* A,       \
                                        const rocblas_int lda, rocblas_int* ipiv,               \
                                        rocblas_int* info)                                      \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri", "-n", n, "--lda", lda);                             \
        return getri_impl<T>(handle, n, strided(A, 0), A != nullptr, lda, ipiv, 0, false,       \
                             strided<T>(nullptr, 0), false, 0, info, 1, true);                  \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_batched(rocblas_handle handle, const rocblas_int n,     \
                                                T* const A[], const rocblas_int lda,            \
                                                rocblas_int* ipiv, const rocblas_stride strideP, \
                                                rocblas_int* info, const rocblas_int batch_count) \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_batched", "-n", n, "--lda", lda, "--strideP", strideP, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, arrayed(A), A != nullptr, lda, ipiv, strideP, false,    \
                             strided<T>(nullptr, 0), false, 0, info, batch_count, true);        \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_strided_batched(                                        \
        rocblas_handle handle, const rocblas_int n, T* A, const rocblas_int lda,                \
        const rocblas_stride strideA, rocblas_int* ipiv, const rocblas_stride strideP,          \
        rocblas_int* info, const rocblas_int batch_count)                                       \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_strided_batched", "-n", n, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, strided(A, strideA), A != nullptr, lda, ipiv, strideP,   \
                             false, strided<T>(nullptr, 0), false, 0, info, batch_count, true); \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt(rocblas_handle handle, const rocblas_int n, T* A,  \
                                             const rocblas_int lda, rocblas_int* info)          \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt", "-n", n, "--lda", lda);                        \
        return getri_impl<T>(handle, n, strided(A, 0), A != nullptr, lda, nullptr, 0, false,    \
                             strided<T>(nullptr, 0), false, 0, info, 1, false);                 \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt_batched(rocblas_handle handle, const rocblas_int n, \
                                                     T* const A[], const rocblas_int lda,       \
                                                     rocblas_int* info,                         \
                                                     const rocblas_int batch_count)             \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt_batched", "-n", n, "--lda", lda, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, arrayed(A), A != nullptr, lda, nullptr, 0, false,       \
                             strided<T>(nullptr, 0), false, 0, info, batch_count, false);       \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt_strided_batched(                                   \
        rocblas_handle handle, const rocblas_int n, T* A, const rocblas_int lda,                \
        const rocblas_stride strideA, rocblas_int* info, const rocblas_int batch_count)         \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt_strided_batched", "-n", n, "--lda", lda, "--strideA", strideA, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, strided(A, strideA), A != nullptr, lda, nullptr, 0,      \
                             false, strided<T>(nullptr, 0), false, 0, info, batch_count, false); \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_outofplace(rocblas_handle handle, const rocblas_int n,  \
                                                   T* A, const rocblas_int lda,                 \
                                                   rocblas_int* ipiv, T* C,                     \
                                                   const rocblas_int ldc, rocblas_int* info)    \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_outofplace", "-n", n, "--lda", lda, "--ldc", ldc);    \
        return getri_impl<T>(handle, n, strided(A, 0), A != nullptr, lda, ipiv, 0, true,       \
                             strided(C, 0), C != nullptr, ldc, info, 1, true);                  \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_outofplace_batched(                                     \
        rocblas_handle handle, const rocblas_int n, T* const A[], const rocblas_int lda,        \
        rocblas_int* ipiv, const rocblas_stride strideP, T* const C[], const rocblas_int ldc,   \
        rocblas_int* info, const rocblas_int batch_count)                                       \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_outofplace_batched", "-n", n, "--lda", lda, "--strideP", strideP, "--ldc", ldc, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, arrayed(A), A != nullptr, lda, ipiv, strideP, true,     \
                             arrayed(C), C != nullptr, ldc, info, batch_count, true);           \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_outofplace_strided_batched(                             \
        rocblas_handle handle, const rocblas_int n, T* A, const rocblas_int lda,                \
        const rocblas_stride strideA, rocblas_int* ipiv, const rocblas_stride strideP, T* C,    \
        const rocblas_int ldc, const rocblas_stride strideC, rocblas_int* info,                 \
        const rocblas_int batch_count)                                                          \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_outofplace_strided_batched", "-n", n, "--lda", lda, "--strideA", strideA, "--strideP", strideP, "--ldc", ldc, "--strideC", strideC, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, strided(A, strideA), A != nullptr, lda, ipiv, strideP,   \
                             true, strided(C, strideC), C != nullptr, ldc, info, batch_count,   \
                             true);                                                             \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt_outofplace(rocblas_handle handle,                  \
                                                        const rocblas_int n, T* A,              \
                                                        const rocblas_int lda, T* C,            \
                                                        const rocblas_int ldc,                  \
                                                        rocblas_int* info)                      \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt_outofplace", "-n", n, "--lda", lda, "--ldc", ldc);\
        return getri_impl<T>(handle, n, strided(A, 0), A != nullptr, lda, nullptr, 0, true,     \
                             strided(C, 0), C != nullptr, ldc, info, 1, false);                 \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt_outofplace_batched(                                \
        rocblas_handle handle, const rocblas_int n, T* const A[], const rocblas_int lda,        \
        T* const C[], const rocblas_int ldc, rocblas_int* info, const rocblas_int batch_count)  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt_outofplace_batched", "-n", n, "--lda", lda, "--ldc", ldc, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, arrayed(A), A != nullptr, lda, nullptr, 0, true,        \
                             arrayed(C), C != nullptr, ldc, info, batch_count, false);          \
    }                                                                                           \
    rocblas_status rocsolver_##P##getri_npvt_outofplace_strided_batched(                        \
        rocblas_handle handle, const rocblas_int n, T* A, const rocblas_int lda,                \
        const rocblas_stride strideA, T* C, const rocblas_int ldc, const rocblas_stride strideC, \
        rocblas_int* info, const rocblas_int batch_count)                                       \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "getri_npvt_outofplace_strided_batched", "-n", n, "--lda", lda, "--strideA", strideA, "--ldc", ldc, "--strideC", strideC, "--batch_count", batch_count);\
        return getri_impl<T>(handle, n, strided(A, strideA), A != nullptr, lda, nullptr, 0,      \
                             true, strided(C, strideC), C != nullptr, ldc, info, batch_count,   \
                             false);                                                            \
    }
RSA_GETRI(s, float)
RSA_GETRI(d, double)

#define RSA_CHOLESKY(P, T)                                                                      \
    rocblas_status rocsolver_##P##potf2(rocblas_handle handle, const rocblas_fill uplo,         \
                                        const rocblas_int n, T* A, const rocblas_int lda,       \
                                        rocblas_int* info)                                      \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potf2", "--uplo", uplo, "-n", n, "--lda", lda);             \
        return potrf_impl<T>(handle, uplo, n, strided(A, 0), A != nullptr, lda, info, 1);       \
    }                                                                                           \
    rocblas_status rocsolver_##P##potf2_batched(rocblas_handle handle, const rocblas_fill uplo, \
                                                const rocblas_int n, T* const A[],              \
                                                const rocblas_int lda, rocblas_int* info,       \
                                                const rocblas_int batch_count)                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potf2_batched", "--uplo", uplo, "-n", n, "--lda", lda, "--batch_count", batch_count);\
        return potrf_impl<T>(handle, uplo, n, arrayed(A), A != nullptr, lda, info, batch_count); \
    }                                                                                           \
    rocblas_status rocsolver_##P##potf2_strided_batched(                                        \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n, T* A,              \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* info,                 \
        const rocblas_int batch_count)                                                          \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potf2_strided_batched", "--uplo", uplo, "-n", n, "--lda", lda, "--strideA", strideA, "--batch_count", batch_count);\
        return potrf_impl<T>(handle, uplo, n, strided(A, strideA), A != nullptr, lda, info,     \
                             batch_count);                                                      \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrf(rocblas_handle handle, const rocblas_fill uplo,         \
                                        const rocblas_int n, T* A, const rocblas_int lda,       \
                                        rocblas_int* info)                                      \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrf", "--uplo", uplo, "-n", n, "--lda", lda);             \
        return potrf_impl<T>(handle, uplo, n, strided(A, 0), A != nullptr, lda, info, 1);       \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrf_batched(rocblas_handle handle, const rocblas_fill uplo, \
                                                const rocblas_int n, T* const A[],              \
                                                const rocblas_int lda, rocblas_int* info,       \
                                                const rocblas_int batch_count)                  \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrf_batched", "--uplo", uplo, "-n", n, "--lda", lda, "--batch_count", batch_count);\
        return potrf_impl<T>(handle, uplo, n, arrayed(A), A != nullptr, lda, info, batch_count); \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrf_strided_batched(                                        \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n, T* A,              \
        const rocblas_int lda, const rocblas_stride strideA, rocblas_int* info,                 \
        const rocblas_int batch_count)                                                          \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrf_strided_batched", "--uplo", uplo, "-n", n, "--lda", lda, "--strideA", strideA, "--batch_count", batch_count);\
        return potrf_impl<T>(handle, uplo, n, strided(A, strideA), A != nullptr, lda, info,     \
                             batch_count);                                                      \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrs(rocblas_handle handle, const rocblas_fill uplo,         \
                                        const rocblas_int n, const rocblas_int nrhs, T* A,      \
                                        const rocblas_int lda, T* B, const rocblas_int ldb)     \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrs", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb);\
        return potrs_impl<T>(handle, uplo, n, nrhs, strided(A, 0), A != nullptr, lda,           \
                             strided(B, 0), B != nullptr, ldb, nullptr, 1, false);              \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrs_batched(                                                \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,                    \
        const rocblas_int nrhs, T* const A[], const rocblas_int lda, T* const B[],              \
        const rocblas_int ldb, const rocblas_int batch_count)                                   \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrs_batched", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb, "--batch_count", batch_count);\
        return potrs_impl<T>(handle, uplo, n, nrhs, arrayed(A), A != nullptr, lda, arrayed(B),  \
                             B != nullptr, ldb, nullptr, batch_count, false);                   \
    }                                                                                           \
    rocblas_status rocsolver_##P##potrs_strided_batched(                                        \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,                    \
        const rocblas_int nrhs, T* A, const rocblas_int lda, const rocblas_stride strideA, T* B, \
        const rocblas_int ldb, const rocblas_stride strideB, const rocblas_int batch_count)     \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "potrs_strided_batched", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideA", strideA, "--ldb", ldb, "--strideB", strideB, "--batch_count", batch_count);\
        return potrs_impl<T>(handle, uplo, n, nrhs, strided(A, strideA), A != nullptr, lda,     \
                             strided(B, strideB), B != nullptr, ldb, nullptr, batch_count,      \
                             false);                                                            \
    }                                                                                           \
    rocblas_status rocsolver_##P##posv(rocblas_handle handle, const rocblas_fill uplo,          \
                                       const rocblas_int n, const rocblas_int nrhs, T* A,       \
                                       const rocblas_int lda, T* B, const rocblas_int ldb,      \
                                       rocblas_int* info)                                       \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "posv", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb);\
        return potrs_impl<T>(handle, uplo, n, nrhs, strided(A, 0), A != nullptr, lda,           \
                             strided(B, 0), B != nullptr, ldb, info, 1, true);                  \
    }                                                                                           \
    rocblas_status rocsolver_##P##posv_batched(                                                 \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,                    \
        const rocblas_int nrhs, T* const A[], const rocblas_int lda, T* const B[],              \
        const rocblas_int ldb, rocblas_int* info, const rocblas_int batch_count)                \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "posv_batched", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--ldb", ldb, "--batch_count", batch_count);\
        return potrs_impl<T>(handle, uplo, n, nrhs, arrayed(A), A != nullptr, lda, arrayed(B),  \
                             B != nullptr, ldb, info, batch_count, true);                       \
    }                                                                                           \
    rocblas_status rocsolver_##P##posv_strided_batched(                                         \
        rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,                    \
        const rocblas_int nrhs, T* A, const rocblas_int lda, const rocblas_stride strideA, T* B, \
        const rocblas_int ldb, const rocblas_stride strideB, rocblas_int* info,                 \
        const rocblas_int batch_count)                                                          \
    {                                                                                           \
        RSA_LOG_TOP(handle, P, "posv_strided_batched", "--uplo", uplo, "-n", n, "--nrhs", nrhs, "--lda", lda, "--strideA", strideA, "--ldb", ldb, "--strideB", strideB, "--batch_count", batch_count);\
        return potrs_impl<T>(handle, uplo, n, nrhs, strided(A, strideA), A != nullptr, lda,     \
                             strided(B, strideB), B != nullptr, ldb, info, batch_count, true);  \
    }
RSA_CHOLESKY(s, float)
RSA_CHOLESKY(d, double)

#define RSA_VBATCHED(P, T)                                                                      \
    rocblas_status rocsolver_amd_##P##getrf_vbatched(                                           \
        rocblas_handle handle, const rocblas_int* n, T* const A[], const rocblas_int* lda,      \
        rocblas_int* const ipiv[], rocblas_int* info, const rocblas_int batch_count)            \
    {                                                                                           \
        return vbatched_impl<T>(handle, 0, rocblas_operation_none, n, nullptr, A, lda, ipiv,    \
                                nullptr, nullptr, info, batch_count);                           \
    }                                                                                           \
    rocblas_status rocsolver_amd_##P##getrs_vbatched(                                           \
        rocblas_handle handle, const rocblas_operation trans, const rocblas_int* n,             \
        const rocblas_int* nrhs, T* const A[], const rocblas_int* lda,                          \
        rocblas_int* const ipiv[], T* const B[], const rocblas_int* ldb, rocblas_int* info,     \
        const rocblas_int batch_count)                                                          \
    {                                                                                           \
        return vbatched_impl<T>(handle, 1, trans, n, nrhs, A, lda, ipiv, B, ldb, info,          \
                                batch_count);                                                   \
    }                                                                                           \
    rocblas_status rocsolver_amd_##P##gesv_vbatched(                                            \
        rocblas_handle handle, const rocblas_int* n, const rocblas_int* nrhs, T* const A[],     \
        const rocblas_int* lda, rocblas_int* const ipiv[], T* const B[],                        \
        const rocblas_int* ldb, rocblas_int* info, const rocblas_int batch_count)               \
    {                                                                                           \
        return vbatched_impl<T>(handle, 2, rocblas_operation_none, n, nrhs, A, lda, ipiv, B,    \
                                ldb, info, batch_count);                                        \
    }
RSA_VBATCHED(s, float)
RSA_VBATCHED(d, double)

rocblas_status rocsolver_amd_dgesv_vbatched_noalloc(rocblas_handle handle, const rocblas_int* n,
                                                   const rocblas_int* nrhs, double* const A[],
                                                   const rocblas_int* lda,
                                                   rocblas_int* const ipiv[], double* const B[],
                                                   const rocblas_int* ldb, rocblas_int* info,
                                                   const rocblas_int batch_count)
{
    return dgesv_vbatched_noalloc_impl(handle, n, nrhs, A, lda, ipiv, B, ldb, info, batch_count);
}

rocblas_status rocsolver_amd_vbatched_set_tuning(const rocblas_int* params, const rocblas_int len)
{
    return vbatched_set_tuning(params, len) ? rocblas_status_success
                                            : rocblas_status_invalid_value;
}

rocblas_status rocsolver_amd_vbatched_get_tuning(rocblas_int* params, const rocblas_int len)
{
    if(!params)
        return rocblas_status_invalid_pointer;
    if(len != VB_TUNING_LEN)
        return rocblas_status_invalid_size;
    const VbTuning t = vbatched_tuning();
    for(int k = 0; k < VB_BUCKETS; ++k)
    {
        params[k] = t.team[k];
        params[VB_BUCKETS + k] = t.pack[k];
    }
    params[2 * VB_BUCKETS] = t.wg_per_cu;
    return rocblas_status_success;
}

#define RSA_VERSION_STRING "3.23.0.gfx950"

rocblas_status rocsolver_get_version_string(char* buf, size_t len)
{
    static constexpr char v[] = RSA_VERSION_STRING;
    if(!buf)
        return rocblas_status_invalid_pointer;
    if(len < sizeof(v))
        return rocblas_status_invalid_size;
    std::memcpy(buf, v, sizeof(v));
    return rocblas_status_success;
}

rocblas_status rocsolver_get_version_string_size(size_t* len)
{
    if(!len)
        return rocblas_status_invalid_pointer;
    *len = std::strlen(RSA_VERSION_STRING) + 1;
    return rocblas_status_success;
}

} // extern "C"
