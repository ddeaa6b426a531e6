// Persistent variable-size batched gesv kernel for float (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(float, VB_GESV)
} // namespace rsa
