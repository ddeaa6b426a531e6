// lu_fused_ss.hip -- instantiates the fused left-looking LU (lu_fused.hpp) for
// float input, float factorisation.
#include "lu_fused.hpp"

namespace rsa
{
template <>
rocblas_status lu_fused<float, float>(hipStream_t stream, int n, Batch<float> AI, int ldi, Batch<float> AO,
                                int ldo, int* ipiv, int64_t strideP, int* info, int bc,
                                Batch<float> WI, const int* skip)
{
    return lu_fused_launch<float, float>(stream, n, AI, ldi, AO, ldo, ipiv, strideP, info, bc,
                                             WI, skip);
}
} // namespace rsa
