// lu_small_gesv_double.hip -- instantiates the register-resident small-matrix gesv for double
// (kernels in lu_small.hpp; one family and type per translation unit so the
// unrolled kernels compile in parallel).
#include "lu_small.hpp"

namespace rsa
{
template <> rocblas_status gesv_small<double>(hipStream_t s, int n, int nrhs, Batch<double> A, int lda, int* ipiv,
                                         int64_t strideP, Batch<double> B, int ldb, int* info, int bc)
{
    return gesv_small_launch<double>(s, n, nrhs, A, lda, ipiv, strideP, B, ldb, info, bc);
}
} // namespace rsa
