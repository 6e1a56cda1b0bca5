// FP32 instantiation of the stencil kernels (mixed-precision path).
#include "stencils_impl.hpp"
template struct Launch<float>;
