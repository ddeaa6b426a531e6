// getrs_fused_ss.hip -- instantiates the one-workgroup-per-system solve
// (getrs_fused.hpp) for a float LU and float right-hand sides.
#include "getrs_fused.hpp"

namespace rsa
{
template <>
rocblas_status getrs_fused<float, float>(hipStream_t stream, bool trans, int n, int nrhs, Batch<float> A,
                                   int lda, const int* ipiv, int64_t strideP, Batch<float> B, int ldb,
                                   Batch<float> X, int ldx, int bc, const int* skip, bool accum)
{
    if(accum)
        return rocblas_status_not_implemented; // only the mixed FP32/FP64 form accumulates
    return getrs_fused_launch<float, float, false>(stream, trans, n, nrhs, A, lda, ipiv, strideP, B, ldb,
                                             X, ldx, bc, skip);
}
} // namespace rsa
