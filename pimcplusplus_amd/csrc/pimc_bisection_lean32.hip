// pimc_bisection_lean32.hip -- instantiates the lean kernel (DavidPA fits on linear grids, no long range), mixed precision:
// every block size, launch mode and number of off-diagonal terms.
#include "pimc_bisection.hpp"

namespace pimc {
PIMC_BISECTION_VARIANT(true, kLeanDavid)
}  // namespace pimc
