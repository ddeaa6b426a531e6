/* rocsolver_amd.h -- C ABI of the MI355X (gfx950) batched LU engine.
 *
 * Drop-in for the batched dense LU factor-and-solve path of rocsolver.h:
 * every function below has the signature, argument meaning, argument-check
 * order, status codes and in/out semantics of the rocSOLVER function of the
 * same name.  Citations are to the reference source tree
 * (library/include/rocsolver/rocsolver-functions.h and the implementation
 * files named).
 *
 * Conventions shared by all functions (as in the reference):
 *  - matrices are column-major, element (i,j) at A[i + j*lda];
 *  - all matrix, pivot and info arrays are caller-owned DEVICE memory;
 *  - _strided_batched: matrix b at A + b*strideA; _batched: A is a device
 *    array of bc device pointers; ipiv is strided by strideP in both forms;
 *  - ipiv is 1-based; info[b] = first zero pivot (1-based) or 0, and a zero
 *    pivot is not an error status;
 *  - work is queued on rocblas_get_stream(handle), asynchronous to the host,
 *    with no host synchronisation inside a call.
 *
 * Extensions with no reference counterpart are prefixed rocsolver_amd_.
 */
#ifndef ROCSOLVER_AMD_H
#define ROCSOLVER_AMD_H

#include <rocblas/rocblas.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ GETRF
 * LU factorisation with partial pivoting, A = P L U.
 * Reference: rocsolver-functions.h:5139/5147 (getrf), 5228/5238 (_batched),
 * 5329/5340 (_strided_batched); implementation
 * library/src/lapack/roclapack_getrf{,_batched,_strided_batched}.cpp. */
rocblas_status rocsolver_sgetrf(rocblas_handle handle, const rocblas_int m, const rocblas_int n,
                                float* A, const rocblas_int lda, rocblas_int* ipiv,
                                rocblas_int* info);
rocblas_status rocsolver_dgetrf(rocblas_handle handle, const rocblas_int m, const rocblas_int n,
                                double* A, const rocblas_int lda, rocblas_int* ipiv,
                                rocblas_int* info);
rocblas_status rocsolver_sgetrf_batched(rocblas_handle handle, const rocblas_int m,
                                        const rocblas_int n, float* const A[],
                                        const rocblas_int lda, rocblas_int* ipiv,
                                        const rocblas_stride strideP, rocblas_int* info,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_dgetrf_batched(rocblas_handle handle, const rocblas_int m,
                                        const rocblas_int n, double* const A[],
                                        const rocblas_int lda, rocblas_int* ipiv,
                                        const rocblas_stride strideP, rocblas_int* info,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_sgetrf_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                const rocblas_int n, float* A,
                                                const rocblas_int lda,
                                                const rocblas_stride strideA, rocblas_int* ipiv,
                                                const rocblas_stride strideP, rocblas_int* info,
                                                const rocblas_int batch_count);
rocblas_status rocsolver_dgetrf_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                const rocblas_int n, double* A,
                                                const rocblas_int lda,
                                                const rocblas_stride strideA, rocblas_int* ipiv,
                                                const rocblas_stride strideP, rocblas_int* info,
                                                const rocblas_int batch_count);

/* Non-pivoting LU, A = L U.  Reference: rocsolver-functions.h:4609/4616
 * (getrf_npvt), 4687/4695 (_batched), 4773/4782 (_strided_batched);
 * roclapack_getrf_strided_batched.cpp:154-208. */
rocblas_status rocsolver_sgetrf_npvt(rocblas_handle handle, const rocblas_int m,
                                     const rocblas_int n, float* A, const rocblas_int lda,
                                     rocblas_int* info);
rocblas_status rocsolver_dgetrf_npvt(rocblas_handle handle, const rocblas_int m,
                                     const rocblas_int n, double* A, const rocblas_int lda,
                                     rocblas_int* info);
rocblas_status rocsolver_sgetrf_npvt_batched(rocblas_handle handle, const rocblas_int m,
                                             const rocblas_int n, float* const A[],
                                             const rocblas_int lda, rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_dgetrf_npvt_batched(rocblas_handle handle, const rocblas_int m,
                                             const rocblas_int n, double* const A[],
                                             const rocblas_int lda, rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_sgetrf_npvt_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                     const rocblas_int n, float* A,
                                                     const rocblas_int lda,
                                                     const rocblas_stride strideA,
                                                     rocblas_int* info,
                                                     const rocblas_int batch_count);
rocblas_status rocsolver_dgetrf_npvt_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                     const rocblas_int n, double* A,
                                                     const rocblas_int lda,
                                                     const rocblas_stride strideA,
                                                     rocblas_int* info,
                                                     const rocblas_int batch_count);

/* Unblocked LU (same results contract as getrf).  Reference:
 * rocsolver-functions.h:4857/4865 (getf2), 4946/4956 (_batched), 5047/5058
 * (_strided_batched); library/src/lapack/roclapack_getf2{,_batched,_strided_batched}.cpp. */
rocblas_status rocsolver_sgetf2(rocblas_handle handle, const rocblas_int m, const rocblas_int n,
                                float* A, const rocblas_int lda, rocblas_int* ipiv,
                                rocblas_int* info);
rocblas_status rocsolver_dgetf2(rocblas_handle handle, const rocblas_int m, const rocblas_int n,
                                double* A, const rocblas_int lda, rocblas_int* ipiv,
                                rocblas_int* info);
rocblas_status rocsolver_sgetf2_batched(rocblas_handle handle, const rocblas_int m,
                                        const rocblas_int n, float* const A[],
                                        const rocblas_int lda, rocblas_int* ipiv,
                                        const rocblas_stride strideP, rocblas_int* info,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_dgetf2_batched(rocblas_handle handle, const rocblas_int m,
                                        const rocblas_int n, double* const A[],
                                        const rocblas_int lda, rocblas_int* ipiv,
                                        const rocblas_stride strideP, rocblas_int* info,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_sgetf2_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                const rocblas_int n, float* A,
                                                const rocblas_int lda,
                                                const rocblas_stride strideA, rocblas_int* ipiv,
                                                const rocblas_stride strideP, rocblas_int* info,
                                                const rocblas_int batch_count);
rocblas_status rocsolver_dgetf2_strided_batched(rocblas_handle handle, const rocblas_int m,
                                                const rocblas_int n, double* A,
                                                const rocblas_int lda,
                                                const rocblas_stride strideA, rocblas_int* ipiv,
                                                const rocblas_stride strideP, rocblas_int* info,
                                                const rocblas_int batch_count);

/* ------------------------------------------------------------------ GETRS
 * Solve op(A) X = B with the LU from getrf; B overwritten by X.
 * Reference: rocsolver-functions.h:8562/8572 (getrs), 8657/8669 (_batched),
 * 8768/8782 (_strided_batched); library/src/lapack/roclapack_getrs.hpp. */
rocblas_status rocsolver_sgetrs(rocblas_handle handle, const rocblas_operation trans,
                                const rocblas_int n, const rocblas_int nrhs, float* A,
                                const rocblas_int lda, const rocblas_int* ipiv, float* B,
                                const rocblas_int ldb);
rocblas_status rocsolver_dgetrs(rocblas_handle handle, const rocblas_operation trans,
                                const rocblas_int n, const rocblas_int nrhs, double* A,
                                const rocblas_int lda, const rocblas_int* ipiv, double* B,
                                const rocblas_int ldb);
rocblas_status rocsolver_sgetrs_batched(rocblas_handle handle, const rocblas_operation trans,
                                        const rocblas_int n, const rocblas_int nrhs,
                                        float* const A[], const rocblas_int lda,
                                        const rocblas_int* ipiv, const rocblas_stride strideP,
                                        float* const B[], const rocblas_int ldb,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_dgetrs_batched(rocblas_handle handle, const rocblas_operation trans,
                                        const rocblas_int n, const rocblas_int nrhs,
                                        double* const A[], const rocblas_int lda,
                                        const rocblas_int* ipiv, const rocblas_stride strideP,
                                        double* const B[], const rocblas_int ldb,
                                        const rocblas_int batch_count);
rocblas_status rocsolver_sgetrs_strided_batched(rocblas_handle handle,
                                                const rocblas_operation trans,
                                                const rocblas_int n, const rocblas_int nrhs,
                                                float* A, const rocblas_int lda,
                                                const rocblas_stride strideA,
                                                const rocblas_int* ipiv,
                                                const rocblas_stride strideP, float* B,
                                                const rocblas_int ldb,
                                                const rocblas_stride strideB,
                                                const rocblas_int batch_count);
rocblas_status rocsolver_dgetrs_strided_batched(rocblas_handle handle,
                                                const rocblas_operation trans,
                                                const rocblas_int n, const rocblas_int nrhs,
                                                double* A, const rocblas_int lda,
                                                const rocblas_stride strideA,
                                                const rocblas_int* ipiv,
                                                const rocblas_stride strideP, double* B,
                                                const rocblas_int ldb,
                                                const rocblas_stride strideB,
                                                const rocblas_int batch_count);

/* ------------------------------------------------------------------- GESV
 * Solve A X = B (getrf + getrs); A overwritten by its LU, B by X, except that
 * B is left unchanged for systems with info != 0.
 * Reference: rocsolver-functions.h:8872/8882 (gesv), 8968/8980 (_batched),
 * 9081/9095 (_strided_batched); library/src/lapack/roclapack_gesv.hpp. */
rocblas_status rocsolver_sgesv(rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs,
                               float* A, const rocblas_int lda, rocblas_int* ipiv, float* B,
                               const rocblas_int ldb, rocblas_int* info);
rocblas_status rocsolver_dgesv(rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs,
                               double* A, const rocblas_int lda, rocblas_int* ipiv, double* B,
                               const rocblas_int ldb, rocblas_int* info);
rocblas_status rocsolver_sgesv_batched(rocblas_handle handle, const rocblas_int n,
                                       const rocblas_int nrhs, float* const A[],
                                       const rocblas_int lda, rocblas_int* ipiv,
                                       const rocblas_stride strideP, float* const B[],
                                       const rocblas_int ldb, rocblas_int* info,
                                       const rocblas_int batch_count);
rocblas_status rocsolver_dgesv_batched(rocblas_handle handle, const rocblas_int n,
                                       const rocblas_int nrhs, double* const A[],
                                       const rocblas_int lda, rocblas_int* ipiv,
                                       const rocblas_stride strideP, double* const B[],
                                       const rocblas_int ldb, rocblas_int* info,
                                       const rocblas_int batch_count);
rocblas_status rocsolver_sgesv_strided_batched(rocblas_handle handle, const rocblas_int n,
                                               const rocblas_int nrhs, float* A,
                                               const rocblas_int lda,
                                               const rocblas_stride strideA, rocblas_int* ipiv,
                                               const rocblas_stride strideP, float* B,
                                               const rocblas_int ldb,
                                               const rocblas_stride strideB, rocblas_int* info,
                                               const rocblas_int batch_count);
rocblas_status rocsolver_dgesv_strided_batched(rocblas_handle handle, const rocblas_int n,
                                               const rocblas_int nrhs, double* A,
                                               const rocblas_int lda,
                                               const rocblas_stride strideA, rocblas_int* ipiv,
                                               const rocblas_stride strideP, double* B,
                                               const rocblas_int ldb,
                                               const rocblas_stride strideB, rocblas_int* info,
                                               const rocblas_int batch_count);

/* ------------------------------------------------------------------ GETRI
 * Inverse from the getrf factors (LU and ipiv) in place, or into C
 * (out of place); the npvt forms use factors made without pivoting (no ipiv).
 * info = first zero diagonal entry of U (1-based); the inverse is then
 * undefined.  The in-place forms take n*n*batch_count elements of handle
 * workspace (size query supported).  Reference declarations
 * (rocsolver-functions.h, s / d): getri :9173/:9180, getri_batched
 * :9245/:9254, getri_strided_batched :9329/:9339, getri_npvt :9402/:9408,
 * getri_npvt_batched :9463/:9470, getri_npvt_strided_batched :9532/:9540,
 * getri_outofplace :21380/:21389, getri_outofplace_batched :21462/:21473,
 * getri_outofplace_strided_batched :21562/:21576, getri_npvt_outofplace
 * :21653/:21661, getri_npvt_outofplace_batched :21725/:21734,
 * getri_npvt_outofplace_strided_batched :21813/:21825; implementation
 * library/src/lapack/roclapack_getri*.{cpp,hpp}. */
rocblas_status rocsolver_sgetri(rocblas_handle handle, const rocblas_int n, float* A,
    const rocblas_int lda, rocblas_int* ipiv, rocblas_int* info);
rocblas_status rocsolver_sgetri_batched(rocblas_handle handle, const rocblas_int n,
    float* const A[], const rocblas_int lda, rocblas_int* ipiv, const rocblas_stride strideP,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_strided_batched(rocblas_handle handle, const rocblas_int n,
    float* A, const rocblas_int lda, const rocblas_stride strideA, rocblas_int* ipiv,
    const rocblas_stride strideP, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_npvt(rocblas_handle handle, const rocblas_int n, float* A,
    const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_sgetri_npvt_batched(rocblas_handle handle, const rocblas_int n,
    float* const A[], const rocblas_int lda, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_npvt_strided_batched(rocblas_handle handle, const rocblas_int n,
    float* A, const rocblas_int lda, const rocblas_stride strideA, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_outofplace(rocblas_handle handle, const rocblas_int n, float* A,
    const rocblas_int lda, rocblas_int* ipiv, float* C, const rocblas_int ldc, rocblas_int* info);
rocblas_status rocsolver_sgetri_outofplace_batched(rocblas_handle handle, const rocblas_int n,
    float* const A[], const rocblas_int lda, rocblas_int* ipiv, const rocblas_stride strideP,
    float* const C[], const rocblas_int ldc, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_outofplace_strided_batched(rocblas_handle handle,
    const rocblas_int n, float* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* ipiv, const rocblas_stride strideP, float* C, const rocblas_int ldc,
    const rocblas_stride strideC, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_npvt_outofplace(rocblas_handle handle, const rocblas_int n,
    float* A, const rocblas_int lda, float* C, const rocblas_int ldc, rocblas_int* info);
rocblas_status rocsolver_sgetri_npvt_outofplace_batched(rocblas_handle handle, const rocblas_int n,
    float* const A[], const rocblas_int lda, float* const C[], const rocblas_int ldc,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sgetri_npvt_outofplace_strided_batched(rocblas_handle handle,
    const rocblas_int n, float* A, const rocblas_int lda, const rocblas_stride strideA, float* C,
    const rocblas_int ldc, const rocblas_stride strideC, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_dgetri(rocblas_handle handle, const rocblas_int n, double* A,
    const rocblas_int lda, rocblas_int* ipiv, rocblas_int* info);
rocblas_status rocsolver_dgetri_batched(rocblas_handle handle, const rocblas_int n,
    double* const A[], const rocblas_int lda, rocblas_int* ipiv, const rocblas_stride strideP,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_strided_batched(rocblas_handle handle, const rocblas_int n,
    double* A, const rocblas_int lda, const rocblas_stride strideA, rocblas_int* ipiv,
    const rocblas_stride strideP, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_npvt(rocblas_handle handle, const rocblas_int n, double* A,
    const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_dgetri_npvt_batched(rocblas_handle handle, const rocblas_int n,
    double* const A[], const rocblas_int lda, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_npvt_strided_batched(rocblas_handle handle, const rocblas_int n,
    double* A, const rocblas_int lda, const rocblas_stride strideA, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_outofplace(rocblas_handle handle, const rocblas_int n, double* A,
    const rocblas_int lda, rocblas_int* ipiv, double* C, const rocblas_int ldc, rocblas_int* info);
rocblas_status rocsolver_dgetri_outofplace_batched(rocblas_handle handle, const rocblas_int n,
    double* const A[], const rocblas_int lda, rocblas_int* ipiv, const rocblas_stride strideP,
    double* const C[], const rocblas_int ldc, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_outofplace_strided_batched(rocblas_handle handle,
    const rocblas_int n, double* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* ipiv, const rocblas_stride strideP, double* C, const rocblas_int ldc,
    const rocblas_stride strideC, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_npvt_outofplace(rocblas_handle handle, const rocblas_int n,
    double* A, const rocblas_int lda, double* C, const rocblas_int ldc, rocblas_int* info);
rocblas_status rocsolver_dgetri_npvt_outofplace_batched(rocblas_handle handle, const rocblas_int n,
    double* const A[], const rocblas_int lda, double* const C[], const rocblas_int ldc,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dgetri_npvt_outofplace_strided_batched(rocblas_handle handle,
    const rocblas_int n, double* A, const rocblas_int lda, const rocblas_stride strideA, double* C,
    const rocblas_int ldc, const rocblas_stride strideC, rocblas_int* info,
    const rocblas_int batch_count);

/* --------------------------------------------------------------- CHOLESKY
 * potf2 / potrf: Cholesky factorisation of the uplo triangle in place (lower:
 * A = L L^T, upper: A = U^T U); info = first non-positive pivot (1-based), at
 * which the factorisation stores the pivot, scales the rest of its column
 * (row) by 0 and goes on, as the reference does.  potrs: solve A X = B with
 * the factor.  posv: potrf + potrs, B unchanged where info != 0.  No
 * workspace.  Reference declarations (rocsolver-functions.h, s / d): potf2
 * :9959/:9966, potf2_batched :10031/:10039, potf2_strided_batched
 * :10111/:10120, potrf :10188/:10195, potrf_batched :10260/:10268,
 * potrf_strided_batched :10340/:10349, potrs :10426/:10435, potrs_batched
 * :10517/:10527, potrs_strided_batched :10620/:10632, posv :10718/:10728,
 * posv_batched :10812/:10823, posv_strided_batched :10918/:10931;
 * implementation library/src/lapack/roclapack_{potf2,potrf,potrs,posv}*. */
rocblas_status rocsolver_spotf2(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* A, const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_spotf2_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* const A[], const rocblas_int lda, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_spotf2_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_spotrf(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* A, const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_spotrf_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* const A[], const rocblas_int lda, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_spotrf_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, float* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_spotrs(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, float* A, const rocblas_int lda, float* B,
    const rocblas_int ldb);
rocblas_status rocsolver_spotrs_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, float* const A[], const rocblas_int lda,
    float* const B[], const rocblas_int ldb, const rocblas_int batch_count);
rocblas_status rocsolver_spotrs_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, float* A, const rocblas_int lda,
    const rocblas_stride strideA, float* B, const rocblas_int ldb, const rocblas_stride strideB,
    const rocblas_int batch_count);
rocblas_status rocsolver_sposv(rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,
    const rocblas_int nrhs, float* A, const rocblas_int lda, float* B, const rocblas_int ldb,
    rocblas_int* info);
rocblas_status rocsolver_sposv_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, float* const A[], const rocblas_int lda,
    float* const B[], const rocblas_int ldb, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_sposv_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, float* A, const rocblas_int lda,
    const rocblas_stride strideA, float* B, const rocblas_int ldb, const rocblas_stride strideB,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dpotf2(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* A, const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_dpotf2_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* const A[], const rocblas_int lda, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_dpotf2_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dpotrf(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* A, const rocblas_int lda, rocblas_int* info);
rocblas_status rocsolver_dpotrf_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* const A[], const rocblas_int lda, rocblas_int* info,
    const rocblas_int batch_count);
rocblas_status rocsolver_dpotrf_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, double* A, const rocblas_int lda, const rocblas_stride strideA,
    rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dpotrs(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, double* A, const rocblas_int lda, double* B,
    const rocblas_int ldb);
rocblas_status rocsolver_dpotrs_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, double* const A[], const rocblas_int lda,
    double* const B[], const rocblas_int ldb, const rocblas_int batch_count);
rocblas_status rocsolver_dpotrs_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, double* A, const rocblas_int lda,
    const rocblas_stride strideA, double* B, const rocblas_int ldb, const rocblas_stride strideB,
    const rocblas_int batch_count);
rocblas_status rocsolver_dposv(rocblas_handle handle, const rocblas_fill uplo, const rocblas_int n,
    const rocblas_int nrhs, double* A, const rocblas_int lda, double* B, const rocblas_int ldb,
    rocblas_int* info);
rocblas_status rocsolver_dposv_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, double* const A[], const rocblas_int lda,
    double* const B[], const rocblas_int ldb, rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_dposv_strided_batched(rocblas_handle handle, const rocblas_fill uplo,
    const rocblas_int n, const rocblas_int nrhs, double* A, const rocblas_int lda,
    const rocblas_stride strideA, double* B, const rocblas_int ldb, const rocblas_stride strideB,
    rocblas_int* info, const rocblas_int batch_count);

/* ------------------------------------------------------------------ LASWP
 * Row interchanges driven by ipiv[k1-1 .. k2-1] (incx = +-1).
 * Reference: rocsolver-functions.h:198/207;
 * library/src/auxiliary/rocauxiliary_laswp.{cpp,hpp}. */
rocblas_status rocsolver_slaswp(rocblas_handle handle, const rocblas_int n, float* A,
                                const rocblas_int lda, const rocblas_int k1, const rocblas_int k2,
                                const rocblas_int* ipiv, const rocblas_int incx);
rocblas_status rocsolver_dlaswp(rocblas_handle handle, const rocblas_int n, double* A,
                                const rocblas_int lda, const rocblas_int k1, const rocblas_int k2,
                                const rocblas_int* ipiv, const rocblas_int incx);

/* ------------------------------------------------- mixed-precision solve
 * No counterpart in the reference (it has no refinement solver); semantics
 * of LAPACK DSGESV applied to every matrix of a strided batch:
 * FP32 LU of A on the matrix cores, FP32 triangular solves, FP64 residual
 * R = B - A X and update, stopping when every column satisfies
 * ||R||_inf <= ||X||_inf ||A||_inf sqrt(n) 2^-53, at most 30 corrections;
 * matrices that do not converge (or whose FP32 copy overflows or is
 * singular) are solved again with an FP64 LU.
 *   A, B   device, read only (A is not overwritten, unlike DSGESV's fallback)
 *   X      device, ldx >= n: the solution
 *   iter   device, per matrix: >= 0 corrections used; -1 FP64 path by
 *          size (below); -2 FP32 overflow; -3 FP32 LU singular; -31 no
 *          convergence (FP64 path used)
 *   info   device, per matrix: first zero pivot of the FP64 LU on the
 *          fallback path, else 0 (X is defined only where info == 0)
 *   berr   device or NULL: per matrix max_c ||R_c||_inf / (||A||_inf ||X_c||_inf)
 *          at convergence (-1 where the FP64 path produced X)
 * Workspace comes from the handle (size query supported): about
 * bc * (8 n^2 + 4 n) bytes.  Refinement runs for n <= 512 with A, B and X
 * each under 2^31 bytes per matrix; otherwise every system is solved by an
 * FP64 LU of a copy of A and iter = -1 (DSGESV: fell back to full precision
 * for implementation-specific reasons), berr = -1.  With
 * nrhs > 8 (> 4 when n > 256) the columns are refined in groups of 8 (4):
 * each group iterates until its own columns pass the test (its solution is
 * the one a call with those columns alone gives), and iter is the largest
 * group count; DSGESV instead refines every column until all pass, so a
 * group that passed early gets no further corrections here. */
rocblas_status rocsolver_amd_dsgesv_strided_batched(
    rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, const double* A,
    const rocblas_int lda, const rocblas_stride strideA, const double* B, const rocblas_int ldb,
    const rocblas_stride strideB, double* X, const rocblas_int ldx, const rocblas_stride strideX,
    rocblas_int* iter, rocblas_int* info, double* berr, const rocblas_int batch_count);

/* Mixed-precision solve with a choice of factorisation (no counterpart in
 * the reference; same arguments, outputs and fallback as
 * rocsolver_amd_dsgesv_strided_batched):
 *   factor 0: FP32 LU of A (identical to rocsolver_amd_dsgesv_strided_batched);
 *   factor 1: A is equilibrated with power-of-two row and column scales R, C,
 *             and R A C is factored in FP32 with the trailing updates on the
 *             matrix cores in FP16 operands (FP32 accumulation);
 *   factor 2: the same with BF16 operands.
 * For factors 1 and 2 each correction solves (R A C) y = R r and adds C y; the
 * residual and the stopping test use A itself in FP64, and entries of A beyond
 * the FP32 range are not a failure, and the right-hand side of each FP32
 * solve is normalised by a power of two (iter = -2 does not occur).  Refinement converges when the half-width factors are
 * accurate enough for the matrix (roughly kappa(A) * 2^-11 < 1 for FP16,
 * kappa(A) * 2^-8 < 1 for BF16); other systems take the FP64 path (iter =
 * -31).  An invalid factor gives rocblas_status_invalid_value. */
rocblas_status rocsolver_amd_dxgesv_strided_batched(
    rocblas_handle handle, const rocblas_int factor, const rocblas_int n, const rocblas_int nrhs,
    const double* A, const rocblas_int lda, const rocblas_stride strideA, const double* B,
    const rocblas_int ldb, const rocblas_stride strideB, double* X, const rocblas_int ldx,
    const rocblas_stride strideX, rocblas_int* iter, rocblas_int* info, double* berr,
    const rocblas_int batch_count);

/* FP32 LU of FP64 matrices (the factorisation step of the mixed-precision
 * solve, usable on its own).  No counterpart in the reference.  A (FP64,
 * read only) is rounded to FP32 and factored with partial pivoting into LU
 * (FP32, ldlu >= n), with getrf's ipiv/info conventions. */
rocblas_status rocsolver_amd_dsgetrf_strided_batched(
    rocblas_handle handle, const rocblas_int n, const double* A, const rocblas_int lda,
    const rocblas_stride strideA, float* LU, const rocblas_int ldlu, const rocblas_stride strideLU,
    rocblas_int* ipiv, const rocblas_stride strideP, rocblas_int* info,
    const rocblas_int batch_count);

/* ---------------------------------------------------------------- adjoint
 * Reverse-mode derivative of X = A^-1 B for every system of a strided batch.
 * No counterpart in the reference.  Inputs: the LU and ipiv of A (from getrf
 * or gesv), the solution X and the upstream gradient dX.  Outputs:
 *   dB = A^-T dX   (transposed getrs with the stored LU and ipiv; dB may alias dX)
 *   dA = -dB X^T
 * All arrays device, column-major; no workspace. */
rocblas_status rocsolver_amd_sgesv_adjoint_strided_batched(
    rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, const float* LU,
    const rocblas_int lda, const rocblas_stride strideA, const rocblas_int* ipiv,
    const rocblas_stride strideP, const float* X, const rocblas_int ldx,
    const rocblas_stride strideX, const float* dX, const rocblas_int lddx,
    const rocblas_stride strideDX, float* dB, const rocblas_int lddb,
    const rocblas_stride strideDB, float* dA, const rocblas_int ldda,
    const rocblas_stride strideDA, const rocblas_int batch_count);
rocblas_status rocsolver_amd_dgesv_adjoint_strided_batched(
    rocblas_handle handle, const rocblas_int n, const rocblas_int nrhs, const double* LU,
    const rocblas_int lda, const rocblas_stride strideA, const rocblas_int* ipiv,
    const rocblas_stride strideP, const double* X, const rocblas_int ldx,
    const rocblas_stride strideX, const double* dX, const rocblas_int lddx,
    const rocblas_stride strideDX, double* dB, const rocblas_int lddb,
    const rocblas_stride strideDB, double* dA, const rocblas_int ldda,
    const rocblas_stride strideDA, const rocblas_int batch_count);

/* ------------------------------------------------------- variable size
 * Batches whose matrices have different orders (0 <= n[b] <= 512).  No
 * counterpart in the reference.  Sizes are device arrays of batch_count
 * entries: n, lda (and nrhs, ldb for getrs/gesv); A, B and ipiv are device
 * arrays of device pointers (ipiv[b] holds n[b] entries).  Per matrix the
 * result equals the fixed-size routine's (getrf / getrs / gesv semantics,
 * 1-based ipiv, info = first zero pivot, singular gesv leaves B unchanged).
 * A matrix with invalid sizes (n < 0, n > 512, lda < max(1,n), nrhs < 0,
 * ldb < max(1,n), lda*n*sizeof(T) >= 2^31) is skipped with info[b] = -1.
 * Runs as two launches with no host synchronisation (plan + one persistent
 * kernel), so it can be captured in a graph once the handle's workspace
 * (size query supported: about 20 bytes per matrix) is allocated.
 * getrs takes an optional info array (NULL allowed): 0, or -1 when skipped. */
rocblas_status rocsolver_amd_sgetrf_vbatched(rocblas_handle handle, const rocblas_int* n,
                                             float* const A[], const rocblas_int* lda,
                                             rocblas_int* const ipiv[], rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_amd_dgetrf_vbatched(rocblas_handle handle, const rocblas_int* n,
                                             double* const A[], const rocblas_int* lda,
                                             rocblas_int* const ipiv[], rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_amd_sgetrs_vbatched(rocblas_handle handle, const rocblas_operation trans,
                                             const rocblas_int* n, const rocblas_int* nrhs,
                                             float* const A[], const rocblas_int* lda,
                                             rocblas_int* const ipiv[], float* const B[],
                                             const rocblas_int* ldb, rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_amd_dgetrs_vbatched(rocblas_handle handle, const rocblas_operation trans,
                                             const rocblas_int* n, const rocblas_int* nrhs,
                                             double* const A[], const rocblas_int* lda,
                                             rocblas_int* const ipiv[], double* const B[],
                                             const rocblas_int* ldb, rocblas_int* info,
                                             const rocblas_int batch_count);
rocblas_status rocsolver_amd_sgesv_vbatched(rocblas_handle handle, const rocblas_int* n,
                                            const rocblas_int* nrhs, float* const A[],
                                            const rocblas_int* lda, rocblas_int* const ipiv[],
                                            float* const B[], const rocblas_int* ldb,
                                            rocblas_int* info, const rocblas_int batch_count);
rocblas_status rocsolver_amd_dgesv_vbatched(rocblas_handle handle, const rocblas_int* n,
                                            const rocblas_int* nrhs, double* const A[],
                                            const rocblas_int* lda, rocblas_int* const ipiv[],
                                            double* const B[], const rocblas_int* ldb,
                                            rocblas_int* info, const rocblas_int batch_count);

/* rocsolver_amd_dgesv_vbatched with no allocation of any kind.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The work queue lives in a static device array of the
 * library (1.25 MiB per GPU, held until the process ends) instead of handle
 * workspace; a size query reports 0 bytes.  batch_count above 65536 returns
 * rocblas_status_invalid_size.  Results are those of
 * rocsolver_amd_dgesv_vbatched.  Every call on one GPU shares that GPU's
 * array: calls that overlap in time on the same GPU (different streams or
 * threads) race and must be serialised by the caller. */
rocblas_status rocsolver_amd_dgesv_vbatched_noalloc(rocblas_handle handle, const rocblas_int* n,
                                                   const rocblas_int* nrhs, double* const A[],
                                                   const rocblas_int* lda,
                                                   rocblas_int* const ipiv[], double* const B[],
                                                   const rocblas_int* ldb, rocblas_int* info,
                                                   const rocblas_int batch_count);

/* Schedule table of the variable-size kernel (process-wide), 13 integers:
 * team[6] (lanes per matrix for orders <= 8, 16, 24, 32, 48, 64: 16, 32 or
 * 64, at least the order), pack[6] (rounds of teams per work item, 1..64),
 * workgroups per CU (1..4).  Set by the autotuner (rocsolver_amd/autotune.py);
 * invalid tables are rejected with rocblas_status_invalid_value. */
rocblas_status rocsolver_amd_vbatched_set_tuning(const rocblas_int* params,
                                                 const rocblas_int len);
rocblas_status rocsolver_amd_vbatched_get_tuning(rocblas_int* params, const rocblas_int len);

/* ---------------------------------------------------------------- logging
 * Multi-level logging of the rocsolver_* entry points: trace trees, bench
 * lines ("./rocsolver-bench -f <function> -r <precision> <arguments>") and
 * profile totals of the internal stages, per the layer-mode bits 0x1 trace,
 * 0x2 bench, 0x4 profile (profile synchronises the handle's stream after
 * every stage).  log_begin reads ROCSOLVER_LAYER and ROCSOLVER_LEVELS and
 * opens ROCSOLVER_LOG_{TRACE,BENCH,PROFILE}_PATH or ROCSOLVER_LOG_PATH
 * (default stderr); every call without an active session, log_begin with
 * one, and log_end with calls still in flight return
 * rocblas_status_internal_error.  Reference: rocsolver-functions.h:66
 * (log_begin), :75 (log_end), :86 (log_set_layer_mode), :98
 * (log_set_max_levels), :108 (log_restore_defaults), :113
 * (log_write_profile), :119 (log_flush_profile); rocsolver-extra-types.h:13
 * (rocblas_layer_mode_flags); library/src/common/rocsolver_logger.cpp. */
#ifndef ROCSOLVER_EXTRA_TYPES_H
typedef uint32_t rocblas_layer_mode_flags;
#endif
rocblas_status rocsolver_log_begin(void);
rocblas_status rocsolver_log_end(void);
rocblas_status rocsolver_log_set_layer_mode(const rocblas_layer_mode_flags layer_mode);
rocblas_status rocsolver_log_set_max_levels(const rocblas_int max_levels);
rocblas_status rocsolver_log_restore_defaults(void);
rocblas_status rocsolver_log_write_profile(void);
rocblas_status rocsolver_log_flush_profile(void);

/* ---------------------------------------------------------------- version
 * Reference: rocsolver-functions.h:34 and :45; library/src/common/buildinfo.cpp. */
rocblas_status rocsolver_get_version_string(char* buf, size_t len);
rocblas_status rocsolver_get_version_string_size(size_t* len);

#ifdef __cplusplus
}
#endif
#endif /* ROCSOLVER_AMD_H */
