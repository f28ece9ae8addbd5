// pimc_bisection_gen32.hip -- instantiates the generic kernel (any fits, long range), mixed precision:
// every block size, launch mode and number of off-diagonal terms.
#include "pimc_bisection.hpp"

namespace pimc {
PIMC_BISECTION_VARIANT(true, kGeneric)
}  // namespace pimc
