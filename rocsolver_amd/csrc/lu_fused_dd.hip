// lu_fused_dd.hip -- instantiates the fused left-looking LU (lu_fused.hpp) for
// double input, double factorisation.
#include "lu_fused.hpp"

namespace rsa
{
template <>
rocblas_status lu_fused<double, double>(hipStream_t stream, int n, Batch<double> AI, int ldi, Batch<double> AO,
                                int ldo, int* ipiv, int64_t strideP, int* info, int bc,
                                Batch<double> WI, const int* skip)
{
    return lu_fused_launch<double, double>(stream, n, AI, ldi, AO, ldo, ipiv, strideP, info, bc,
                                             WI, skip);
}
} // namespace rsa
