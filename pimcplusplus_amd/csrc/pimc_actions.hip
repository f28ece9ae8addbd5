// pimc_actions.hip -- kernels behind the SingleAction entry points, the fit
// evaluation, DistDisp and the accept/reject/join bookkeeping.
#include "pimc_kernels.hpp"
#include "pimc_launch.hpp"

namespace pimc {

// --------------------------------------------------------------------------
// PairActionFitClass::U for a batch of points (tables read from global memory)
// --------------------------------------------------------------------------
template <int NT, bool MIXED>
__global__ void __launch_bounds__(kBlock) k_fit_eval_u(const FitDev* __restrict__ fits, int fit, int level,
                                                       int64_t n, const double* __restrict__ q,
                                                       const double* __restrict__ z,
                                                       const double* __restrict__ s2, double* __restrict__ out) {
  const FitDev F = fits[fit];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int lev = level + F.taupos;
  if (F.kind == 1)
    out[i] = MIXED ? pair_u_bc_mixed(F, qgrid_of(F), level, F.i32_lev[level], q[i], s2[i])
                   : pair_u_bc(F, qgrid_of(F), level, q[i], s2[i]);
  else if (MIXED)
    out[i] = pair_u<NT>(F, level_table32(F, lev), q[i], z[i], s2[i]);
  else
    out[i] = pair_u<NT>(F, level_table64(F, lev), q[i], z[i], s2[i]);
}

template <bool MIXED>
static void launch_fit_eval_t(int nt, const LaunchFitEval& a, hipStream_t s) {
  dim3 grid((unsigned)((a.n + kBlock - 1) / kBlock));
  switch (nt) {
    case 0: k_fit_eval_u<0, MIXED><<<grid, kBlock, 0, s>>>(a.fits, a.fit, a.level, a.n, a.q, a.z, a.s2, a.out); break;
    case 1: k_fit_eval_u<1, MIXED><<<grid, kBlock, 0, s>>>(a.fits, a.fit, a.level, a.n, a.q, a.z, a.s2, a.out); break;
    case 2: k_fit_eval_u<2, MIXED><<<grid, kBlock, 0, s>>>(a.fits, a.fit, a.level, a.n, a.q, a.z, a.s2, a.out); break;
    default: k_fit_eval_u<3, MIXED><<<grid, kBlock, 0, s>>>(a.fits, a.fit, a.level, a.n, a.q, a.z, a.s2, a.out); break;
  }
}
void launch_fit_eval(int num_terms, bool mixed, const LaunchFitEval& a, hipStream_t s) {
  if (a.n <= 0) return;
  if (mixed)
    launch_fit_eval_t<true>(num_terms, a, s);
  else
    launch_fit_eval_t<false>(num_terms, a, s);
}

// --------------------------------------------------------------------------
// ShortRangeClass::SingleAction for a batch of requests.
// grid.x: partner tiles of kBlock particles times nsplit link ranges, grid.y: requests.
// Each thread owns one partner and runs the moved-particle and link loops of the reference
// for it over its block's range of the links (nsplit > 1 shortens the dependent walk of a
// call with few requests); partial sums land in partials[req][tile * nsplit + range] and a
// second kernel adds them in that order.
// --------------------------------------------------------------------------
template <int NT, bool MIXED>
__global__ void __launch_bounds__(kBlock) k_short_range(const ReplicaDev* __restrict__ reps,
                                                        const FitDev* __restrict__ fits,
                                                        const ReqDev* __restrict__ reqs, int ntiles, int nsplit,
                                                        double* __restrict__ partials) {
  __shared__ double scratch_raw[kWaves];
  PIMC_LDS double* scratch = (PIMC_LDS double*)scratch_raw;
  const ReqDev Q = reqs[blockIdx.y];
  const PIMC_GLOBAL int* qp = (const PIMC_GLOBAL int*)reqs[blockIdx.y].ptcls;   // indexed at run time: not from Q
  const ReplicaDev* RD = reps + Q.replica;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int tile = blockIdx.x / nsplit, range = blockIdx.x % nsplit;
  const int j = tile * blockDim.x + threadIdx.x;
  const int skip = 1 << Q.level;
  const int nlinks = (Q.slice2 - Q.slice1 + skip - 1) / skip;
  const int first = Q.slice1 + (int)((long long)range * nlinks / nsplit) * skip;
  const int last = Q.slice1 + (int)((long long)(range + 1) * nlinks / nsplit) * skip;   // this range's links end here
  Kahan acc;
  if (j < R.N) {
    const int sp2 = [&] {
      int s = 0;
      while (s + 1 < R.ns && j >= S.first[s + 1]) s++;
      return s;
    }();
    for (int i1 = 0; i1 < Q.nptcls; i1++) {
      const int p1 = qp[i1];
      // DoPtcl: partner j is skipped if it is p1 or a mover handled before p1
      bool skipj = false;
      for (int t = 0; t <= i1; t++) skipj |= (qp[t] == j);
      if (skipj) continue;
      int sp1 = 0;
      while (sp1 + 1 < R.ns && p1 >= S.first[sp1 + 1]) sp1++;
      const FitDev F = fits[S.fit[sp1 * R.ns + sp2]];
      const int lev = Q.level + F.taupos;
      const UlongView ul = ulong_for(R, F, sp1 * R.ns + sp2, Q.level);
      double ma[3], mb[3], pa[3], pb[3];
      load_bead(R, Q.mode, first, p1, ma);
      load_bead(R, Q.mode, first, j, pa);
      for (int slice = first; slice < min(last, Q.slice2); slice += skip) {
        load_bead(R, Q.mode, slice + skip, p1, mb);
        load_bead(R, Q.mode, slice + skip, j, pb);
        double u;
        if (F.kind == 1 && MIXED)
          u = pair_link_u<NT>(R, F, BcView32{{qgrid_of(F), Q.level, F.i32_lev[Q.level]}}, ma, mb, pa, pb, ul);
        else if (F.kind == 1)
          u = pair_link_u<NT>(R, F, BcView{qgrid_of(F), Q.level, 0}, ma, mb, pa, pb, ul);
        else if (MIXED)
          u = pair_link_u<NT>(R, F, level_table32(F, lev), ma, mb, pa, pb, ul);
        else
          u = pair_link_u<NT>(R, F, level_table64(F, lev), ma, mb, pa, pb, ul);
        if (MIXED)
          acc.add(u);
        else
          acc.s += u;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          ma[d] = mb[d];
          pa[d] = pb[d];
        }
      }
    }
  }
  double t = block_sum(acc.value(), scratch);
  if (threadIdx.x == 0) partials[(size_t)blockIdx.y * ntiles * nsplit + blockIdx.x] = t;
}

__global__ void k_sum_tiles(const double* __restrict__ partials, int ntiles, int nreq, double* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreq) return;
  double s = 0.0;
  for (int t = 0; t < ntiles; t++) s += partials[(size_t)r * ntiles + t];
  out[r] = s;
}

template <bool MIXED>
static void launch_sr_t(int nt, const LaunchShortRange& a, hipStream_t s) {
  dim3 grid(a.ntiles * a.nsplit, a.nreq);
#define PIMC_SR(NT_) k_short_range<NT_, MIXED><<<grid, kBlock, 0, s>>>(a.reps, a.fits, a.reqs, a.ntiles, a.nsplit, a.partials)
  switch (nt) {
    case 0: PIMC_SR(0); break;
    case 1: PIMC_SR(1); break;
    case 2: PIMC_SR(2); break;
    default: PIMC_SR(3); break;
  }
#undef PIMC_SR
}
// With one block per request (one tile, one link range) the block sums are the results: one
// launch.
void launch_short_range(int num_terms, bool mixed, const LaunchShortRange& a, hipStream_t s) {
  if (a.nreq <= 0) return;
  const int parts = a.ntiles * a.nsplit;
  LaunchShortRange b = a;
  if (parts == 1) b.partials = a.out;
  if (mixed)
    launch_sr_t<true>(num_terms, b, s);
  else
    launch_sr_t<false>(num_terms, b, s);
  if (parts > 1) k_sum_tiles<<<(a.nreq + 255) / 256, 256, 0, s>>>(a.partials, parts, a.nreq, a.out);
}

// --------------------------------------------------------------------------
// KineticClass::SingleAction, one wave per request: lane l takes links l, l + 64, ..., and the
// wave adds the lanes' sums (the links' loads are in flight together instead of one link after
// another)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_kinetic(const ReplicaDev* __restrict__ reps, const ReqDev* __restrict__ reqs,
                                                int nreq, double* __restrict__ out) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  if (r >= nreq) return;
  const ReqDev Q = reqs[r];
  const PIMC_GLOBAL int* qp = (const PIMC_GLOBAL int*)reqs[r].ptcls;   // indexed at run time: not from Q
  const RepCore& R = reps[Q.replica].core;
  const RepSpecies& S = reps[Q.replica].sp;
  const int skip = 1 << Q.level;
  const double levelTau = R.tau * (1 << Q.level);
  double TotalK = 0.0;
  for (int pi = 0; pi < Q.nptcls; pi++) {
    int ptcl = qp[pi];
    int sp = 0;
    while (sp + 1 < R.ns && ptcl >= S.first[sp + 1]) sp++;
    double lambda = S.lambda[sp];
    if (lambda != 0) {
      double FourLambdaTauInv = 1.0 / (4.0 * lambda * levelTau);
      for (int slice = Q.slice1 + lane * skip; slice < Q.slice2; slice += 64 * skip) {
        double a[3], b[3];
        load_bead(R, Q.mode, slice, ptcl, a);
        load_bead(R, Q.mode, slice + skip, ptcl, b);
        TotalK += kinetic_link(R, a, b, FourLambdaTauInv);
      }
    }
  }
  TotalK = wave_sum(TotalK);
  if (lane == 0) out[r] = TotalK;
}
void launch_kinetic(const ReplicaDev* reps, const ReqDev* reqs, int nreq, double* out, hipStream_t s) {
  if (nreq <= 0) return;
  k_kinetic<<<nreq, 64, 0, s>>>(reps, reqs, nreq, out);
}

// --------------------------------------------------------------------------
// PathClass::DistDisp for index quadruples
// --------------------------------------------------------------------------
__global__ void k_dist_disp(const ReplicaDev* __restrict__ reps, int rep, int mode, int64_t n,
                            const int* __restrict__ idx, double* __restrict__ out, int* __restrict__ images) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const RepCore& R = reps[rep].core;
  const int* q = idx + 4 * i;
  double a1[3], a2[3], b1[3], b2[3];
  load_bead(R, mode, q[0], q[2], a1);
  load_bead(R, mode, q[0], q[3], a2);
  load_bead(R, mode, q[1], q[2], b1);
  load_bead(R, mode, q[1], q[3], b2);
  double da[3], db[3], nn[3], mm[3];
  for (int d = 0; d < 3; d++)
    dist_disp_1(a2[d] - a1[d], b2[d] - b1[d], R.box[d], R.boxinv[d], R.isper[d], da[d], db[d], nn[d], mm[d]);
  double* o = out + 8 * i;
  o[0] = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
  o[1] = sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]);
  for (int d = 0; d < 3; d++) {
    o[2 + d] = da[d];
    o[5 + d] = db[d];
    images[6 * i + d] = (int)nn[d];
    images[6 * i + 3 + d] = (int)mm[d];
  }
}
void launch_dist_disp(const ReplicaDev* reps, int rep, int mode, int64_t n, const int* idx, double* out,
                      int* images, hipStream_t s) {
  if (n <= 0) return;
  k_dist_disp<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(reps, rep, mode, n, idx, out, images);
}

// --------------------------------------------------------------------------
// PathClass::AcceptCopy / RejectCopy (PathClass.cc:797-899) for up to
// kMaxActive particles; one thread per (particle, slice).
// --------------------------------------------------------------------------
__global__ void k_accept_reject(const ReplicaDev* __restrict__ reps, int rep, int slice1, int slice2,
                                ReqDev plist, int accept) {
  const RepCore& R = reps[rep].core;
  const int ns = slice2 - slice1 + 1;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= plist.nptcls * ns) return;
  const int p = plist.ptcls[t / ns];
  const int slice = slice1 + t % ns;
  double v[3];
  if (accept) {
    load_bead(R, 1, slice, p, v);
    for (int d = 0; d < 3; d++) v[d] = put_in_box_fast_1(v[d], R.box[d]);
    store_bead(R, 1, slice, p, v);
    store_bead(R, 0, slice, p, v);
    if (t % ns == 0) R.perm0[p] = R.perm1[p];
  } else {
    load_bead(R, 0, slice, p, v);
    store_bead(R, 1, slice, p, v);
    if (t % ns == 0) R.perm1[p] = R.perm0[p];
  }
}
void launch_accept_reject(const ReplicaDev* reps, int rep, int slice1, int slice2, const int* ptcls, int n,
                          bool accept, hipStream_t s) {
  ReqDev pl{};
  pl.nptcls = n;
  for (int i = 0; i < n; i++) pl.ptcls[i] = ptcls[i];
  int total = n * (slice2 - slice1 + 1);
  if (total <= 0) return;
  k_accept_reject<<<(total + 255) / 256, 256, 0, s>>>(reps, rep, slice1, slice2, pl, accept ? 1 : 0);
}

// --------------------------------------------------------------------------
// PathClass::MoveJoin (PathClass.cc:727-793), closed paths.  Step 1 writes OLD
// from NEW through the NEW permutation, step 2 copies OLD back to NEW.
// --------------------------------------------------------------------------
__global__ void k_move_join_1(const ReplicaDev* __restrict__ reps, int rep, int lo, int hi, int forward) {
  const RepCore& R = reps[rep].core;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nsl = hi - lo + 1;
  if (t >= nsl * R.N) return;
  int slice = lo + t / R.N, p = t % R.N;
  double v[3];
  if (forward) {
    load_bead(R, 1, slice, R.perm1[p], v);
    store_bead(R, 0, slice, p, v);
  } else {
    load_bead(R, 1, slice, p, v);
    store_bead(R, 0, slice, R.perm1[p], v);
  }
}
__global__ void k_move_join_2(const ReplicaDev* __restrict__ reps, int rep, int lo, int hi) {
  const RepCore& R = reps[rep].core;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int nsl = hi - lo + 1;
  if (t >= nsl * R.N) return;
  int slice = lo + t / R.N, p = t % R.N;
  double v[3];
  load_bead(R, 0, slice, p, v);
  store_bead(R, 1, slice, p, v);
}
void launch_move_join(const ReplicaDev* reps, int rep, int N, int old_join, int new_join, hipStream_t s) {
  int lo, hi, fwd;
  if (new_join > old_join) {
    lo = old_join + 1; hi = new_join; fwd = 1;
  } else if (old_join > new_join) {
    lo = new_join + 1; hi = old_join; fwd = 0;
  } else {
    return;
  }
  int total = (hi - lo + 1) * N;
  k_move_join_1<<<(total + 255) / 256, 256, 0, s>>>(reps, rep, lo, hi, fwd);
  k_move_join_2<<<(total + 255) / 256, 256, 0, s>>>(reps, rep, lo, hi);
}

// --------------------------------------------------------------------------
// Path.SetPos of one particle over a slice range, one thread per (slice, dimension).
// --------------------------------------------------------------------------
__global__ void k_set_beads(const ReplicaDev* __restrict__ reps, int rep, int mode, int ptcl, int slice1, int nsl,
                            const double* __restrict__ xyz) {
  const RepCore& R = reps[rep].core;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nsl) return;
  const int slice = slice1 + t / 3, d = t % 3;
  const size_t at = ((size_t)slice * 3 + d) * R.N + ptcl;
  if (mode != 1) R.path0[at] = xyz[t];
  if (mode != 0) R.path1[at] = xyz[t];
}
void launch_set_beads(const ReplicaDev* reps, int rep, int mode, int ptcl, int slice1, int slice2, const double* xyz,
                      hipStream_t s) {
  const int nsl = slice2 - slice1 + 1;
  k_set_beads<<<(3 * nsl + 255) / 256, 256, 0, s>>>(reps, rep, mode, ptcl, slice1, nsl, xyz);
}

// --------------------------------------------------------------------------
// PathClass::PutInBox (PathClass.h:897-921) or PutInBoxFast (PathClass.cc:1389-1407) of
// n displacement vectors [n][3], one thread per component.
// --------------------------------------------------------------------------
__global__ void k_put_in_box(const ReplicaDev* __restrict__ reps, int rep, int fast, int64_t n, double* v) {
  const RepCore& R = reps[rep].core;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int d = (int)(t % 3);
  v[t] = fast ? put_in_box_fast_1(v[t], R.box[d]) : put_in_box_1(v[t], R.box[d], R.boxinv[d], R.isper[d]);
}
void launch_put_in_box(const ReplicaDev* reps, int rep, int fast, int64_t n, double* v, hipStream_t s) {
  const int64_t total = 3 * n;
  if (total <= 0) return;
  k_put_in_box<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(reps, rep, fast, n, v);
}

// --------------------------------------------------------------------------
// PermuteTableClass::ConstructHTable (PermuteTableClass.cc:122-150): one thread
// per (i, j) of the species' particles; displacements through PutInBox, dot
// products summed over x, y, z in order.
// --------------------------------------------------------------------------
__global__ void k_permute_htable(const ReplicaDev* __restrict__ reps, int rep, int mode, int first, int n, int slice1,
                                 int slice2, double four_lambda_beta_inv, double log_eps, float* __restrict__ out) {
  const RepCore& R = reps[rep].core;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  double a[3], b[3], c[3];
  load_bead(R, mode, slice1, first + i, a);
  load_bead(R, mode, slice2, first + i, b);
  load_bead(R, mode, slice2, first + j, c);
  double dii = 0.0, dij = 0.0;
  for (int d = 0; d < 3; d++) {
#pragma clang fp contract(off)
    const double u = put_in_box_1(b[d] - a[d], R.box[d], R.boxinv[d], R.isper[d]);
    const double w = put_in_box_1(c[d] - a[d], R.box[d], R.boxinv[d], R.isper[d]);
    dii += u * u;
    dij += w * w;
  }
  const double to_exp = (-dij + dii) * four_lambda_beta_inv;
  out[t] = to_exp < log_eps ? 0.0f : (float)exp(to_exp);
}
void launch_permute_htable(const ReplicaDev* reps, int rep, int mode, int first, int n, int slice1, int slice2,
                           double four_lambda_beta_inv, double log_eps, float* out, hipStream_t s) {
  const long long total = (long long)n * n;
  if (total <= 0) return;
  k_permute_htable<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(reps, rep, mode, first, n, slice1, slice2,
                                                                  four_lambda_beta_inv, log_eps, out);
}

// --------------------------------------------------------------------------
// CycleClass::Apply (PermuteTableClass.cc:172-187) on the NEW copy at one slice:
// a rotation of at most four beads and permutation entries, in the reference's
// order, by one thread.
// --------------------------------------------------------------------------
__global__ void k_cycle_apply(const ReplicaDev* __restrict__ reps, int rep, int slice, ReqDev cyc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const RepCore& R = reps[rep].core;
  const int len = cyc.nptcls;
  double temp[3];
  load_bead(R, 1, slice, cyc.ptcls[0], temp);
  const int temp_perm = R.perm1[cyc.ptcls[0]];
  for (int i = 0; i < len - 1; i++) {
    double v[3];
    load_bead(R, 1, slice, cyc.ptcls[i + 1], v);
    store_bead(R, 1, slice, cyc.ptcls[i], v);
    R.perm1[cyc.ptcls[i]] = R.perm1[cyc.ptcls[i + 1]];
  }
  store_bead(R, 1, slice, cyc.ptcls[len - 1], temp);
  R.perm1[cyc.ptcls[len - 1]] = temp_perm;
}
void launch_cycle_apply(const ReplicaDev* reps, int rep, int slice, const int* ptcls, int n, hipStream_t s) {
  ReqDev c{};
  c.nptcls = n;
  for (int i = 0; i < n; i++) c.ptcls[i] = ptcls[i];
  k_cycle_apply<<<1, 64, 0, s>>>(reps, rep, slice, c);
}

}  // namespace pimc
