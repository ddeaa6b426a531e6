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
// when TI == TC).  Batch entries with skip[b] != 0 are not touched.
constexpr int FUSED_MAX = 512;

// WI: workspace for the inverted unit-lower diagonal blocks of the factor,
// lu_fused_inv_elems(n) elements of TC per matrix (Batch addressing).
inline int64_t lu_fused_inv_elems(int n)
{
    return int64_t((n + 31) / 32) * 1024;
}
template <typename TC>
inline size_t lu_fused_workspace_bytes(int n, int bc)
{
    return sizeof(TC) * size_t(lu_fused_inv_elems(n)) * size_t(bc);
}

template <typename TI, typename TC>
rocblas_status lu_fused(hipStream_t stream, int n, Batch<TI> AI, int ldi, Batch<TC> AO, int ldo,
                        int* ipiv, int64_t strideP, int* info, int bc, Batch<TC> WI,
                        const int* skip = nullptr);

// FP32 LU of diag(rs) A diag(cs) (A FP64; the scales of matrix b at
// rs + b * strideS, cs + b * strideS) whose trailing updates run on the matrix
// cores with FP16 (opk = 1) or BF16 (opk = 2) operands and FP32 accumulation.
// Both return the factors with every 32x32 diagonal block inverted in place
// (strictly lower part inv(L_tt), upper part with the diagonal inv(U_tt);
// a partial last block as diag(block, I)), the form the refinement's
// triangular solves use; the plain FP32 LU of an FP64 matrix otherwise.
rocblas_status lu_fused_lowp(hipStream_t stream, int opk, int n, Batch<double> AI, int ldi,
                             Batch<float> AO, int ldo, int* ipiv, int64_t strideP, int* info,
                             int bc, const double* rs, const double* cs, int64_t strideS);
rocblas_status lu_fused_invdiag(hipStream_t stream, int n, Batch<double> AI, int ldi,
                                Batch<float> AO, int ldo, int* ipiv, int64_t strideP, int* info,
                                int bc);

// One-workgroup-per-system solve with the LU of getrf for SMALL_MAX < n <=
// FUSED_MAX, arithmetic in TL; reads the right-hand sides from B (type TB),
// writes (or with accum, adds) the solution to X (may alias B).  Entries with
// skip[b] != 0 are not touched.
template <typename TL, typename TB>
rocblas_status getrs_fused(hipStream_t stream, bool trans, int n, int nrhs, Batch<TL> A, int lda,
                           const int* ipiv, int64_t strideP, Batch<TB> B, int ldb, Batch<TB> X,
                           int ldx, int bc, const int* skip, bool accum);

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

// ----------------------------------------------- mixed-precision refinement
// FP32 LU + FP64 refinement with LAPACK dsgesv semantics per matrix, for
// SMALL_MAX < n <= FUSED_MAX.  `work` must hold dsgesv_workspace_bytes.
// Three launches, no host synchronisation (graph-capturable).
// factor: 0 = FP32 LU of A (LAPACK dsgesv); 1 / 2 = FP32 LU of the
// equilibrated matrix with FP16 / BF16 matrix-core operands in the updates.
size_t dsgesv_workspace_bytes(int n, int nrhs, int bc, int itermax);
rocblas_status dsgesv_refine(hipStream_t stream, int factor, int n, int nrhs, Batch<double> A,
                             int lda, Batch<double> B, int ldb, Batch<double> X, int ldx,
                             int* iter, int* info, double* berr, int bc, int itermax, void* work);

// ------------------------------------------------------------------ adjoint
// dB = A^-T dX (transposed solve with the LU and ipiv of A), dA = -dB X^T.
template <typename T>
rocblas_status gesv_adjoint(hipStream_t stream, int n, int nrhs, Batch<T> LU, int lda,
                            const int* ipiv, int64_t strideP, Batch<T> X, int ldx, Batch<T> dX,
                            int lddx, Batch<T> dB, int lddb, Batch<T> dA, int ldda, int bc);

// --------------------------------------------------- variable-size batches
// Orders 0 .. FUSED_MAX mixed in one batch; per-matrix sizes in device
// arrays.  Six size buckets for n <= 64 (orders 8, 16, 24, 32, 48, 64).
constexpr int VB_BUCKETS = 6;

// Schedule parameters of the persistent kernel (chosen by the autotuner):
// team[k]: lanes per matrix in bucket k (16, 32 or 64, at least the bucket's
// order); pack[k]: rounds of teams per work item (1..64); wg_per_cu:
// persistent workgroups launched per CU (1..4).
struct VbTuning
{
    int team[VB_BUCKETS];
    int pack[VB_BUCKETS];
    int wg_per_cu;
};
constexpr int VB_TUNING_LEN = 2 * VB_BUCKETS + 1;

VbTuning vbatched_tuning(); // current process-wide table
bool vbatched_set_tuning(const int* params, int len);
// Workspace of the variable-size routines: the order list (one int per
// matrix), the work items (16 bytes each, at most one per matrix plus one
// partial pack per bucket) and the two queue counters, each 256-byte aligned.
constexpr size_t vbatched_workspace_bytes(int bc)
{
    return (sizeof(int) * size_t(bc) + 255) / 256 * 256
           + (16 * (size_t(bc) + VB_BUCKETS) + 255) / 256 * 256 + 256;
}
int device_cu_count();

// Largest batch of the entry point that allocates nothing
// (rocsolver_amd_dgesv_vbatched_noalloc): its work queue lives in a static
// device array of vbatched_workspace_bytes(VB_NOALLOC_MAX_BATCH) bytes
// (1.25 MiB) per GPU.
constexpr int VB_NOALLOC_MAX_BATCH = 65536;
// Device address of that array on the current device (looked up once per
// device), or null if the lookup failed.
void* vbatched_noalloc_scratch();

// op: 0 getrf, 1 getrs, 2 gesv.  info may be null for getrs.
template <typename T>
rocblas_status vbatched(hipStream_t stream, int op, bool trans, const int* n, const int* nrhs,
                        T* const* A, const int* lda, int* const* ipiv, T* const* B,
                        const int* ldb, int* info, int bc, void* work);

// ------------------------------------------------------------------- getri
// X = inv(A) from its LU factors (LU, ipiv; pivot = false: no interchanges):
// info = first zero diagonal entry of U, X = I, then A X = I solved in place
// in X.  X must not alias LU.
template <typename T>
rocblas_status getri_solve(hipStream_t stream, int n, Batch<T> LU, int ldlu, const int* ipiv,
                           int64_t strideP, Batch<T> X, int ldx, int* info, int bc, bool pivot);
template <typename T>
rocblas_status copy_square(hipStream_t stream, int n, Batch<T> S, int lds, Batch<T> D, int ldd,
                           int bc);

// ---------------------------------------------------------------- Cholesky
// In-place factorisation of the lower (up = false: A = L L^T) or upper
// (A = U^T U) triangle; info = first non-positive pivot (1-based) or 0.
template <typename T>
rocblas_status potrf(hipStream_t stream, bool up, int n, Batch<T> A, int lda, int* info, int bc);
// Solve A X = B with that factor, B overwritten by X; systems with
// skip[b] != 0 are left untouched.
template <typename T>
rocblas_status potrs(hipStream_t stream, bool up, int n, int nrhs, Batch<T> A, int lda,
                     Batch<T> B, int ldb, int bc, const int* skip);

// --------------------------------------------------------------- utilities
rocblas_status reset_info(hipStream_t stream, int* info, int bc);

template <typename T>
rocblas_status laswp(hipStream_t stream, int ncols, Batch<T> A, int lda, int k1, int k2,
                     const int* ipiv, int64_t strideP, int incx, int bc, const int* skip_info);

} // namespace rsa
