// pimc_bisection.hip -- launches of the GPU-resident bisection chain.
//
// The kernel templates are in pimc_bisection.hpp; the four (precision, variant)
// instantiations are compiled in pimc_bisection_{lean,gen}{32,64}.hip.  This file
// dispatches a launch to them and holds the small launch-reset kernels.
#include "pimc_bisection.hpp"

namespace pimc {

PIMC_BISECTION_EXTERN(true, kLeanDavid)
PIMC_BISECTION_EXTERN(false, kLeanDavid)
PIMC_BISECTION_EXTERN(true, kGeneric)
PIMC_BISECTION_EXTERN(false, kGeneric)

// Clears the work-queue counter and the cluster barrier counters of a launch
// (a kernel, so that a captured sweep holds no memset node).
__global__ void k_launch_reset(int* queue, int* bar, int nbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (queue && i == 0) *queue = 0;
  if (bar && i < nbar) bar[i] = 0;
}
__global__ void k_sweep_advance(uint32_t* first_move, int num_moves) {
  if (threadIdx.x == 0) *first_move += (uint32_t)num_moves;
}

hipError_t launch_bisection(int num_terms, bool mixed, bool lean, const LaunchBisection& a, int block, int grid,
                            hipStream_t s) {
  if (a.nrep <= 0 || a.num_moves <= 0) return hipSuccess;
  if (bisection_lds_bytes(a) > 64 * 1024) return hipErrorInvalidValue;
  if (block != 256 && block != 512) return hipErrorInvalidValue;
  if (a.link_split != 1 && a.link_split != 2 && a.link_split != 4) return hipErrorInvalidValue;
  if (a.cluster > 1) {
    // every block of the grid must be resident at once: the caller checks the grid
    // against the resident capacity; the barrier's bounded wait is the backstop
    if (a.queue || !a.bar || !a.part || !a.abort_flag || a.cluster > kMaxCluster) return hipErrorInvalidValue;
    grid = a.nrep * a.cluster;
    k_launch_reset<<<(a.nrep + 255) / 256, 256, 0, s>>>(nullptr, a.bar, a.nrep);
  } else if (a.queue) {
    k_launch_reset<<<1, 64, 0, s>>>(a.queue, nullptr, 0);
  } else {
    grid = a.nrep;
  }
  if (lean)
    return mixed ? launch_bis_t<true, kLeanDavid>(num_terms, a, block, grid, s)
                 : launch_bis_t<false, kLeanDavid>(num_terms, a, block, grid, s);
  return mixed ? launch_bis_t<true, kGeneric>(num_terms, a, block, grid, s)
               : launch_bis_t<false, kGeneric>(num_terms, a, block, grid, s);
}

hipError_t launch_sweep_advance(uint32_t* first_move, int num_moves, hipStream_t s) {
  k_sweep_advance<<<1, 64, 0, s>>>(first_move, num_moves);
  return hipGetLastError();
}

size_t bisection_lds_bytes_for(const LaunchBisection& a) { return bisection_lds_bytes(a); }

int bisection_blocks_per_cu(int num_terms, bool mixed, bool lean, int block, size_t lds, int mode) {
  int n = 0;
  const int nt = num_terms > 3 ? 3 : num_terms;
  const hipError_t e = lean ? (mixed ? occupancy_t<true, kLeanDavid>(n, nt, block, lds, mode)
                                      : occupancy_t<false, kLeanDavid>(n, nt, block, lds, mode))
                             : (mixed ? occupancy_t<true, kGeneric>(n, nt, block, lds, mode)
                                      : occupancy_t<false, kGeneric>(n, nt, block, lds, mode));
  return e == hipSuccess ? n : 0;
}

}  // namespace pimc
