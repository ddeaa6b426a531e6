// lu_fused_ds.hip -- instantiates the fused left-looking LU (lu_fused.hpp) for
// double input, float factorisation: plain factors (getrf contract), or with
// the diagonal blocks returned inverted (the mixed-precision solvers).
#include "lu_fused.hpp"

namespace rsa
{
template <>
rocblas_status lu_fused<double, float>(hipStream_t stream, int n, Batch<double> AI, int ldi, Batch<float> AO,
                                int ldo, int* ipiv, int64_t strideP, int* info, int bc,
                                Batch<float> WI, const int* skip)
{
    return lu_fused_launch<double, float>(stream, n, AI, ldi, AO, ldo, ipiv, strideP, info, bc,
                                             WI, skip);
}

rocblas_status lu_fused_invdiag(hipStream_t stream, int n, Batch<double> AI, int ldi,
                                Batch<float> AO, int ldo, int* ipiv, int64_t strideP, int* info,
                                int bc)
{
    return lu_fused_launch<double, float, MFMA_NATIVE, 2>(stream, n, AI, ldi, AO, ldo, ipiv,
                                                          strideP, info, bc, Batch<float>{},
                                                          nullptr);
}
} // namespace rsa
