// pimc_launch.hpp -- host-side launch wrappers of the kernels.
#pragma once

#include <vector>

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
  double* partials;  // [nreq][ntiles * nsplit]
  double* out;       // [nreq]
  int nsplit = 1;    // blocks sharing a tile's partners, each with a contiguous range of the links
};
void launch_short_range(int num_terms, bool mixed, const LaunchShortRange& a, hipStream_t s);

void launch_kinetic(const ReplicaDev* reps, const ReqDev* reqs, int nreq, double* out, hipStream_t s);
void launch_dist_disp(const ReplicaDev* reps, int rep, int mode, int64_t n, const int* idx, double* out,
                      int* images, hipStream_t s);
void launch_accept_reject(const ReplicaDev* reps, int rep, int slice1, int slice2, const int* ptcls, int n,
                          bool accept, hipStream_t s);
void launch_move_join(const ReplicaDev* reps, int rep, int N, int old_join, int new_join, hipStream_t s);
// Path.SetPos of one particle over slices [slice1, slice2] from xyz [(slice2-slice1+1)][3]
// (device memory), into copy `mode` (2: both).
void launch_set_beads(const ReplicaDev* reps, int rep, int mode, int ptcl, int slice1, int slice2, const double* xyz,
                      hipStream_t s);
void launch_put_in_box(const ReplicaDev* reps, int rep, int fast, int64_t n, double* v, hipStream_t s);
void launch_permute_htable(const ReplicaDev* reps, int rep, int mode, int first, int n, int slice1, int slice2,
                           double four_lambda_beta_inv, double log_eps, float* out, hipStream_t s);
void launch_cycle_apply(const ReplicaDev* reps, int rep, int slice, const int* ptcls, int n, hipStream_t s);
void launch_rho_all(const ReplicaDev* reps, int rep, int nslices, int ns, int nk, int stride, hipStream_t s);
// rho_k of copy `mode` at slices slice1 + i skip (i < nsl), recomputed from the positions,
// into out[nsl][ns][nk] (PathClass::CalcRho_ks_Fast, PathClass.cc:616-647).
void launch_rho_slices(const ReplicaDev* reps, int rep, int mode, int slice1, int skip, int nsl, int ns, int nk,
                       int stride, d2v* out, hipStream_t s);
// LongRangeClass::SingleAction requests: k blocks of each request in parallel, requests of
// distinct replicas in one launch (in order otherwise); partials is [nreq][kblocks],
// kblocks = long_range_kblocks(the largest nk of the requests' replicas).
int long_range_kblocks(int max_nk);
// The launches of a long-range call: requests [r0, r1) on distinct replicas, the largest
// species count among them and whether any is a NEWMODE request.  nspecies[r] and modes[r]:
// the species count of request r's replica and its mode.
struct LrGroup {
  int r0, r1, nsp;
  bool upd;
};
std::vector<LrGroup> long_range_groups(int nreq, const int* replicas, const int* nspecies, const int* modes);
void launch_long_range(bool mixed, const ReplicaDev* reps, const ReqDev* reqs, int nreq,
                       const std::vector<LrGroup>& groups, int kblocks, double* partials, double* out,
                       hipStream_t s);
// Whether the phase rows of one slice's moved beads (NEW and OLD copies, FP64) fit the
// request kernel's LDS for a k set of `stride` entries per row.
bool long_range_phase_fits(int stride, int nptcls);

// d_dBeta of the actions over links [slice1, slice2) at `level` (slices slice1, slice1 +
// 2^level, .., slice2) of copy `mode`, as partial sums per block: pairs {du, v} (the
// potential over the same slices, ends weighted 1/2), kinetic, long range per slice
// {homo+hetero, background, k0, and the potential's three}.  A zero block count skips
// that kernel.  The long-range kernel reads rho_k of those slices from lr_rho
// ([slice][species][k], recomputed from the positions).
hipError_t launch_energy(int num_terms, const ReplicaDev* reps, const FitDev* fits, int rep, int mode, int slice1,
                         int slice2, int level, int pair_blocks, int kin_blocks, int lr_slices, const d2v* lr_rho,
                         double* pair_partials, double* kin_partials, double* lr_partials, hipStream_t s);
hipError_t launch_pair_correlation(const ReplicaDev* reps, int rep, int mode, int f1, int n1, int f2, int n2,
                                   int same, double start, double deltainv, double end, int nbins,
                                   unsigned long long* hist, long long nitems, int blocks, hipStream_t s);

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
  int lr_chunk;               // slices whose phase tables are staged together (lds_lr_bytes holds them)
  int link_split;             // threads per partner in the short-range loop: 1, 2 or 4
  const int* rep_pos;         // [nrep] caller index of list entry i (records, stats), or nullptr
  int* queue;                 // work-queue counter of a persistent launch, or nullptr
  const uint32_t* first_move_dev;  // first move number in device memory, or nullptr
  // several blocks per replica (cluster > 1; not with a queue): grid = nrep * cluster
  int cluster;
  int* bar;                   // [nrep] barrier counters, cleared by the launch
  double* part;               // [nrep][2][cluster][4] partial stage sums
  int* abort_flag;            // set when a barrier wait timed out (sticky until the host clears it)
  uint64_t wait_ticks;        // bound of one barrier wait, in device wall-clock ticks
  // TABLE permutation moves (k_bisection_table): per list entry, the cycle weights and cut-off
  const struct PermuteDev* permute;
  size_t lds_htable_bytes;    // the HTable of the largest moved species (n^2 floats)
  void* permute_scratch;      // permute_grid regions of cycle-table scratch
  int permute_cap;            // table entries a region holds
  int permute_grid;           // blocks of a TABLE launch (each takes replicas in turn)
  size_t permute_lds_limit;   // LDS of a TABLE block; what the chain leaves holds cycle tables
};

// TablePermuteStageClass settings of one replica of a TABLE launch.
struct PermuteDev {
  double gamma[4];
  double epsilon;
};
constexpr int kMaxCluster = 32;
// Returns hipSuccess or the launch error.  lean: every pair fit is a DavidPA
// fit on a LINEAR grid and no replica evaluates the long-range action.
// block: 256 or 512 threads; grid: blocks of a persistent launch (a.queue set);
// a cluster launch runs nrep * cluster blocks.
hipError_t launch_bisection(int num_terms, bool mixed, bool lean, const LaunchBisection& a, int block, int grid,
                            hipStream_t s);
hipError_t launch_sweep_advance(uint32_t* first_move, int num_moves, hipStream_t s);
// The chain with PermuteType TABLE (a.permute set): one block of 256 threads per replica.
hipError_t launch_bisection_table(int num_terms, bool mixed, bool lean, const LaunchBisection& a, hipStream_t s);
size_t bisection_table_lds_bytes(const LaunchBisection& a);
// LDS a TABLE block may take so that the CU holds as many blocks as the kernel's registers allow.
size_t bisection_table_lds_limit(size_t lds_per_cu);
// Resident TABLE blocks per CU at `lds` bytes of LDS each (0 on error).
int bisection_table_blocks_per_cu(int num_terms, bool mixed, bool lean, size_t lds);
// Device bytes of the cycle-table scratch of a TABLE launch of `grid` blocks.
size_t bisection_table_scratch_bytes(int cap, int grid);
size_t bisection_lds_bytes_for(const LaunchBisection& a);
// mode: 0 one block per replica, 1 persistent, 2 cluster.
int bisection_blocks_per_cu(int num_terms, bool mixed, bool lean, int block, size_t lds, int mode);

}  // namespace pimc
