// Persistent variable-size batched getrf kernel for float (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(float, VB_GETRF)
} // namespace rsa
