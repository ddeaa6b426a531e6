// lu_fused_lowp.hip -- instantiates the fused left-looking LU (lu_fused.hpp)
// for FP64 input scaled by row/column factors, FP32 factors, FP16 or BF16
// matrix-core operands in the trailing updates; diagonal blocks returned
// inverted (the refinement's triangular solves use them).
#include "lu_fused.hpp"

namespace rsa
{
rocblas_status lu_fused_lowp(hipStream_t stream, int opk, int n, Batch<double> AI, int ldi,
                             Batch<float> AO, int ldo, int* ipiv, int64_t strideP, int* info,
                             int bc, const double* rs, const double* cs, int64_t strideS)
{
    if(opk == MFMA_F16)
        return lu_fused_launch<double, float, MFMA_F16, 2>(stream, n, AI, ldi, AO, ldo, ipiv,
                                                        strideP, info, bc, Batch<float>{}, nullptr, rs, cs,
                                                        strideS);
    if(opk == MFMA_BF16)
        return lu_fused_launch<double, float, MFMA_BF16, 2>(stream, n, AI, ldi, AO, ldo, ipiv,
                                                         strideP, info, bc, Batch<float>{}, nullptr, rs, cs,
                                                         strideS);
    return rocblas_status_invalid_value;
}
} // namespace rsa
