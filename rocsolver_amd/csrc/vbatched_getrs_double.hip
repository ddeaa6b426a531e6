// Persistent variable-size batched getrs kernel for double (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(double, VB_GETRS)
} // namespace rsa
