// pimc_bisection_lean64.hip -- instantiates the lean kernel (DavidPA fits on linear grids, no long range), FP64:
// every block size, launch mode and number of off-diagonal terms.
#include "pimc_bisection.hpp"

namespace pimc {
PIMC_BISECTION_VARIANT(false, kLeanDavid)
}  // namespace pimc
