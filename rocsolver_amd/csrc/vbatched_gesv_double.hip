// Persistent variable-size batched gesv kernel for double (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(double, VB_GESV)
} // namespace rsa
