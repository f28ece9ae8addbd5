// pimc_launch.hpp -- host-side launch wrappers of the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pimc_gpu.h"

namespace pimc {

struct FitDev;
struct ReplicaDev;
struct ReqDev;

struct LaunchFitEval {
  const FitDev* fits;
  int fit, level;
  int64_t n;
  const double *q, *z, *s2;
  double* out;
};
void launch_fit_eval(int num_terms, bool mixed, const LaunchFitEval& a, hipStream_t s);

struct LaunchShortRange {
  const ReplicaDev* reps;
  const FitDev* fits;
  const ReqDev* reqs;
  int nreq, ntiles;
  double* partials;  // [nreq][ntiles]
  double* out;       // [nreq]
};
void launch_short_range(int num_terms, bool mixed, const LaunchShortRange& a, hipStream_t s);

void launch_kinetic(const ReplicaDev* reps, const ReqDev* reqs, int nreq, double* out, hipStream_t s);
void launch_dist_disp(const ReplicaDev* reps, int rep, int mode, int64_t n, const int* idx, double* out,
                      int* images, hipStream_t s);
void launch_accept_reject(const ReplicaDev* reps, int rep, int slice1, int slice2, const int* ptcls, int n,
                          bool accept, hipStream_t s);
void launch_move_join(const ReplicaDev* reps, int rep, int N, int old_join, int new_join, hipStream_t s);
void launch_rho_all(const ReplicaDev* reps, int rep, int nslices, int ns, int nk, int stride, hipStream_t s);
void launch_long_range(bool mixed, const ReplicaDev* reps, const ReqDev* reqs, int nreq, const int* strides,
                       double* out, hipStream_t s);

// The GPU-resident bisection chain: one block per replica, num_moves moves each.
struct LaunchBisection {
  const ReplicaDev* reps;
  const FitDev* fits;
  const int* rep_ids;       // [nrep]
  int nrep;
  int num_levels, lowest_level, species;
  uint64_t seed;
  uint32_t first_move;
  int num_moves;
  pimc_move_record* records;  // [nrep][num_moves] or nullptr
  int64_t* stats;             // [nrep][5]: this launch's counters per block
  int64_t* acc;               // [replica id][5]: counters accumulated over launches
  size_t lds_table_bytes;     // max staged level size over the fits used
  size_t lds_lr_bytes;        // phase tables of the long-range stage term (0 without long range)
};
// Returns hipSuccess or the launch error.  lean: every pair fit is a DavidPA
// fit on a LINEAR grid and no replica evaluates the long-range action.
hipError_t launch_bisection(int num_terms, bool mixed, bool lean, const LaunchBisection& a, hipStream_t s);

}  // namespace pimc
