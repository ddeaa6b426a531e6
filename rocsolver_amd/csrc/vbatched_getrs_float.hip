// Persistent variable-size batched getrs kernel for float (vbatched.hpp).
#include "vbatched.hpp"

namespace rsa
{
RSA_VB_INSTANTIATE(float, VB_GETRS)
} // namespace rsa
