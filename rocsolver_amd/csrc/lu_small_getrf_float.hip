// lu_small_getrf_float.hip -- instantiates the register-resident small-matrix getrf for float
// (kernels in lu_small.hpp; one family and type per translation unit so the
// unrolled kernels compile in parallel).
#include "lu_small.hpp"

namespace rsa
{
template <> rocblas_status getrf_small<float>(hipStream_t s, int m, int n, Batch<float> A, int lda, int* ipiv,
                                          int64_t strideP, int* info, int bc, bool pivot)
{
    return getrf_small_launch<float>(s, m, n, A, lda, ipiv, strideP, info, bc, pivot);
}
} // namespace rsa
