// getrs_fused_sd.hip -- instantiates the one-workgroup-per-system solve
// (getrs_fused.hpp) for a float LU and double right-hand sides.
#include "getrs_fused.hpp"

namespace rsa
{
template <>
rocblas_status getrs_fused<float, double>(hipStream_t stream, bool trans, int n, int nrhs, Batch<float> A,
                                   int lda, const int* ipiv, int64_t strideP, Batch<double> B, int ldb,
                                   Batch<double> X, int ldx, int bc, const int* skip, bool accum)
{
    return accum ? getrs_fused_launch<float, double, true>(stream, trans, n, nrhs, A, lda, ipiv, strideP, B, ldb,
                                                  X, ldx, bc, skip)
                 : getrs_fused_launch<float, double, false>(stream, trans, n, nrhs, A, lda, ipiv, strideP, B,
                                                   ldb, X, ldx, bc, skip);
}
} // namespace rsa
