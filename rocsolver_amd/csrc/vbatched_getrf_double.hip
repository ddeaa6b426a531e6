// Persistent variable-size batched getrf kernel for double (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(double, VB_GETRF)
} // namespace rsa
