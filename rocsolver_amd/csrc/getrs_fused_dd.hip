// getrs_fused_dd.hip -- instantiates the one-workgroup-per-system solve
// (getrs_fused.hpp) for a double LU and double right-hand sides.
#include "getrs_fused.hpp"

namespace rsa
{
template <>
rocblas_status getrs_fused<double, double>(hipStream_t stream, bool trans, int n, int nrhs, Batch<double> A,
                                   int lda, const int* ipiv, int64_t strideP, Batch<double> B, int ldb,
                                   Batch<double> X, int ldx, int bc, const int* skip, bool accum)
{
    if(accum)
        return rocblas_status_not_implemented; // only the mixed FP32/FP64 form accumulates
    return getrs_fused_launch<double, double, false>(stream, trans, n, nrhs, A, lda, ipiv, strideP, B, ldb,
                                             X, ldx, bc, skip);
}
} // namespace rsa
