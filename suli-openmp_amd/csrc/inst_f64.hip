// FP64 instantiation of the stencil kernels.  Built with -ffp-contract=off:
// every expression is evaluated in the reference's order with IEEE double
// operations, so the FP64 path matches the reference to rounding of the
// transcendental functions.
#include "stencils_impl.hpp"
template struct Launch<double>;
