// internal.hpp -- host-side interface between the C ABI layer (api.cpp) and
// the kernel drivers (*.hip).  Every driver is stream-ordered, allocates
// nothing, and returns a rocblas_status.
#pragma once

#include "common.hpp"

namespace rsa
{

// Largest m, n handled by the register-resident small kernels.
constexpr int SMALL_MAX = 64;

// ---------------------------------------------------------------- small path
template <typename T>
rocblas_status getrf_small(hipStream_t stream, int m, int n, Batch<T> A, int lda, int* ipiv,
                           int64_t strideP, int* info, int bc, bool pivot);

template <typename T>
rocblas_status gesv_small(hipStream_t stream, int n, int nrhs, Batch<T> A, int lda, int* ipiv,
                          int64_t strideP, Batch<T> B, int ldb, int* info, int bc);

template <typename T>
rocblas_status getrs_small(hipStream_t stream, bool trans, int n, int nrhs, Batch<T> A, int lda,
                           const int* ipiv, int64_t strideP, Batch<T> B, int ldb, int bc,
                           bool pivot);

// ---------------------------------------------------------- fused LU path
// Square LU for SMALL_MAX < n <= FUSED_MAX, one workgroup per matrix.  AI is
// read (type TI), the factors are written to AO (type TC; AO may alias AI
// when TI == TC).
constexpr int FUSED_MAX = 256;

template <typename TI, typename TC>
rocblas_status lu_fused(hipStream_t stream, int n, Batch<TI> AI, int ldi, Batch<TC> AO, int ldo,
                        int* ipiv, int64_t strideP, int* info, int bc);

// -------------------------------------------------------------- blocked path
template <typename T>
rocblas_status getrf_blocked(hipStream_t stream, int m, int n, Batch<T> A, int lda, int* ipiv,
                             int64_t strideP, int* info, int bc, bool pivot);

// skip_info: when non-null, batch entries with skip_info[b] != 0 are left
// untouched (gesv leaves B unchanged for singular systems).
template <typename T>
rocblas_status getrs_blocked(hipStream_t stream, bool trans, int n, int nrhs, Batch<T> A, int lda,
                             const int* ipiv, int64_t strideP, Batch<T> B, int ldb, int bc,
                             bool pivot, const int* skip_info);

// --------------------------------------------------------------- utilities
rocblas_status reset_info(hipStream_t stream, int* info, int bc);

template <typename T>
rocblas_status laswp(hipStream_t stream, int ncols, Batch<T> A, int lda, int k1, int k2,
                     const int* ipiv, int64_t strideP, int incx, int bc, const int* skip_info);

} // namespace rsa
