// lu_small_getrs_float.hip -- instantiates the register-resident small-matrix getrs for float
// (kernels in lu_small.hpp; one family and type per translation unit so the
// unrolled kernels compile in parallel).
#include "lu_small.hpp"

namespace rsa
{
template <> rocblas_status getrs_small<float>(hipStream_t s, bool trans, int n, int nrhs, Batch<float> A,
                                          int lda, const int* ipiv, int64_t strideP, Batch<float> B,
                                          int ldb, int bc, bool pivot)
{
    return getrs_small_launch<float>(s, trans, n, nrhs, A, lda, ipiv, strideP, B, ldb, bc, pivot);
}
} // namespace rsa
