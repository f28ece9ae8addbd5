// pimc_bisection_gen64.hip -- instantiates the generic kernel (any fits, long range), FP64:
// every block size, launch mode and number of off-diagonal terms.
#include "pimc_bisection.hpp"

namespace pimc {
PIMC_BISECTION_VARIANT(false, kGeneric)
}  // namespace pimc
