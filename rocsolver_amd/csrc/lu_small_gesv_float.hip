// lu_small_gesv_float.hip -- instantiates the register-resident small-matrix gesv for float
// (kernels in lu_small.hpp; one family and type per translation unit so the
// unrolled kernels compile in parallel).
#include "lu_small.hpp"

namespace rsa
{
template <> rocblas_status gesv_small<float>(hipStream_t s, int n, int nrhs, Batch<float> A, int lda, int* ipiv,
                                         int64_t strideP, Batch<float> B, int ldb, int* info, int bc)
{
    return gesv_small_launch<float>(s, n, nrhs, A, lda, ipiv, strideP, B, ldb, info, bc);
}
} // namespace rsa
