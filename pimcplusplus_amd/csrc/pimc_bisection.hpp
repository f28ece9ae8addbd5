// pimc_bisection.hpp -- the GPU-resident bisection chain (kernel templates).
//
// One block (4 waves) owns one replica and runs `num_moves` bisection moves
// (BisectionBlockClass::MakeMove with NoPermuteStage, StepsPerBlock = 1) in
// order, so the Metropolis chain never leaves the device:
//   ChooseTimeSlices, ChooseParticle            BisectionBlock.cc:363-374, NoPermuteClass.cc:61-68
//   per level L-1 .. LowestLevel:
//     BisectionStageClass::Sample                BisectionStage.cc:128-289
//     StageAction(OLD), StageAction(NEW)         StageClass.cc:124-136 (Kinetic + ShortRange)
//     Metropolis test                            StageClass.cc:69-72
//   AcceptCopy / RejectCopy                      PathClass.cc:797-899
// The moved particle's 2^L+1 beads (OLD and NEW) live in LDS; the level's
// spline table is staged in LDS once per stage; partner beads stream from HBM
// slice-major (one row per dimension, coalesced across the block's threads).
// The path copies OLD and NEW are equal outside a move, so partners are read
// from OLD and accepted beads are written to both.
#pragma once

#include "pimc_kernels.hpp"
#include "pimc_launch.hpp"

namespace pimc {

// One link of one partner (ShortRangeClass.cc:153-167) in the slice walk of
// sr_partners.  On entry dA, rA, uA, lA hold the partner's imaged displacement,
// distance, u00 and Ulong at the link's first slice; on return, those of its
// second slice (nearest image), which is also the next link's first slice.
// DistDisp images the second slice consistently with the first; that image is
// the nearest one unless the partner sits near half a box, where the second
// slice's terms are evaluated again for this link only.
template <int NT, bool MIXED, bool BC, bool LIN, typename Table>
__device__ inline double sr_link(const RepCore& R, const FitDev& F, const Table& T, const BcView& Q,
                                 const UlongView& ul, const double pbox[3], const double hsafe[3],
                                 double urmin, const PIMC_LDS double* m, const double p[3], double dA[3], double& rA,
                                 double& uA, double& lA) {
  double dn[3], raw[3], e[3];
  bool same = true;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    raw[d] = p[d] - m[d];
    {
#pragma clang fp contract(off)
      const double n = -rint(raw[d] * R.boxinv[d]);
      dn[d] = MIXED ? fma(n, pbox[d], raw[d]) : raw[d] + n * pbox[d];   // FP64: DistDisp's rounding
    }
    // DistDisp's consistent image m = nearbyint((dA - raw)/L) equals the nearest
    // image n whenever |dA - dn| < L/2 (hsafe keeps a margin far above rounding)
    e[d] = dA[d] - dn[d];
    same = same && (fabs(e[d]) < hsafe[d]);
  }
  const double n2 = dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2];
  const double rn = MIXED ? sqrt_seeded(n2) : sqrt_dist(n2);
  double un = 0.0, ln = 0.0;
  if constexpr (!BC) un = slice_u00<LIN>(F, T, rn);
  if (ul.t64) ln = ulong_at<MIXED>(R, ul, rn);
  double db[3] = {dn[0], dn[1], dn[2]};
  double rb = rn, ub = un, lb = ln;
  if (!same) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma clang fp contract(off)
      const double mm = rint((dA[d] - raw[d]) * R.boxinv[d]);
      db[d] = MIXED ? fma(mm, pbox[d], raw[d]) : raw[d] + mm * pbox[d];
      e[d] = dA[d] - db[d];
    }
    const double b2 = db[0] * db[0] + db[1] * db[1] + db[2] * db[2];
    rb = MIXED ? sqrt_seeded(b2) : sqrt_dist(b2);
    if constexpr (!BC) ub = slice_u00<LIN>(F, T, rb);
    if (ul.t64) lb = ulong_at<MIXED>(R, ul, rb);
  }
  const double s2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  const double q = 0.5 * (rA + rb);
  const double z = rA - rb;
  double U;
  if constexpr (BC) {
    U = MIXED ? pair_u_bc_mixed(F, Q.q, Q.level, Q.i32, q, s2) : pair_u_bc(F, Q.q, Q.level, q, s2);
  } else {
    // rA < rmin || rb < rmin, as two integer compares: the bit patterns of non-negative doubles
    // (and NaN, above every finite one) order as the values do
    const long long rmin_bits = __double_as_longlong(F.start);
    const bool core = (__double_as_longlong(rA) < rmin_bits) | (__double_as_longlong(rb) < rmin_bits);
    U = core ? urmin : 0.5 * (uA + ub) + link_offdiag<NT, LIN>(F, T, q, z, s2);
  }
  if (ul.t64) U -= 0.5 * (lA + lb);
#pragma unroll
  for (int d = 0; d < 3; d++) dA[d] = dn[d];
  rA = rn;
  uA = un;
  lA = ln;
  return U;
}

// The moved particles a partner loop leaves out: one (the chain without permutations) or the
// particles of a permutation cycle.
struct SkipOne {
  int p;
  __device__ inline bool has(int j) const { return j == p; }
};
struct SkipSet {
  int p[kMaxActive];
  int n;
  __device__ inline bool has(int j) const {
    bool in = false;
#pragma unroll
    for (int i = 0; i < kMaxActive; i++) in = in || (i < n && j == p[i]);
    return in;
  }
};

// ShortRange partner loop of one stage (ShortRangeClass.cc:149-169) for partners
// [j0, j1) other than the moved ones: each thread walks one partner's links in order, OLD and NEW copies
// of the moved particle side by side, two links per iteration with the
// partner's beads loaded two links ahead.  `ls` threads share a partner, each
// taking a contiguous range of K/ls links (at least two; K is a power of two):
// the block covers blockDim/ls partners per pass.  Consecutive lanes take
// consecutive partners, so bead loads coalesce.
template <int NT, bool MIXED, bool BC, bool LIN, typename Table, typename Skip = SkipOne>
__device__ inline void sr_partners(const RepCore& R, const FitDev& F, const Table& T, const BcView& Q,
                                   const UlongView ul, int j0, int j1, Skip skip, int slice1, int K, int step, int ls,
                                   const PIMC_LDS double* mold, const PIMC_LDS double* mnew,
                                   StageSums<MIXED>& acc) {
  const size_t row = (size_t)3 * R.N * step;     // distance between consecutive link slices
  const double pbox[3] = {R.isper[0] * R.box[0], R.isper[1] * R.box[1], R.isper[2] * R.box[2]};
  const double hsafe[3] = {0.4999999 * R.box[0], 0.4999999 * R.box[1], 0.4999999 * R.box[2]};
  double urmin = 0.0;
  if constexpr (!BC) urmin = u00_rmin(T);
  auto load = [&](const PIMC_GLOBAL double* q, double v[3]) {
    v[0] = q[0];
    v[1] = q[R.N];
    v[2] = q[2 * R.N];
  };
  auto link2 = [&](int k, const double p[3], double dO[3], double& rO, double& uO, double& lO, double dN[3],
                   double& rN, double& uN, double& lN) {
    const int at = 3 * (k + 1) * step;
    const double uo = sr_link<NT, MIXED, BC, LIN>(R, F, T, Q, ul, pbox, hsafe, urmin, mold + at, p, dO, rO, uO, lO);
    const double un = sr_link<NT, MIXED, BC, LIN>(R, F, T, Q, ul, pbox, hsafe, urmin, mnew + at, p, dN, rN, uN, lN);
    acc.add(uo, un);
  };
  const int ppass = (int)blockDim.x / ls;             // partners per pass: a multiple of the wave size
  // this thread's link range; whole waves share one, so it is wave-uniform (scalar)
  const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / ppass);
  const int split = (K / 2 >= ls) ? ls : (K / 2);     // ranges in use: each >= 2 links
  if (sub >= split) return;
  const int kper = K / split;
  const int kbeg = sub * kper, kend = kbeg + kper;
  for (int j = j0 + (int)threadIdx.x % ppass; j < j1; j += ppass) {
    if (skip.has(j)) continue;
    const PIMC_GLOBAL double* pj = (const PIMC_GLOBAL double*)R.path0 + (size_t)(slice1 + kbeg * step) * 3 * R.N + j;
    double p0[3], p1[3], p2[3];
    load(pj, p0);
    load(pj + row, p1);
    load(pj + 2 * row, p2);
    const PIMC_LDS double* mo0 = mold + 3 * kbeg * step;
    const PIMC_LDS double* mn0 = mnew + 3 * kbeg * step;
    double dO[3], dN[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma clang fp contract(off)
      const double ao = p0[d] - mo0[d], an = p0[d] - mn0[d];
      if constexpr (MIXED) {
        dO[d] = fma(-rint(ao * R.boxinv[d]), pbox[d], ao);
        dN[d] = fma(-rint(an * R.boxinv[d]), pbox[d], an);
      } else {
        dO[d] = ao + (-rint(ao * R.boxinv[d])) * pbox[d];
        dN[d] = an + (-rint(an * R.boxinv[d])) * pbox[d];
      }
    }
    const double o2 = dO[0] * dO[0] + dO[1] * dO[1] + dO[2] * dO[2];
    const double w2 = dN[0] * dN[0] + dN[1] * dN[1] + dN[2] * dN[2];
    double rO = MIXED ? sqrt_seeded(o2) : sqrt_dist(o2);
    double rN = MIXED ? sqrt_seeded(w2) : sqrt_dist(w2);
    double uO = 0.0, uN = 0.0, lO = 0.0, lN = 0.0;
    if constexpr (!BC) {
      uO = slice_u00<LIN>(F, T, rO);
      uN = slice_u00<LIN>(F, T, rN);
    }
    if (ul.t64) {
      lO = ulong_at<MIXED>(R, ul, rO);
      lN = ulong_at<MIXED>(R, ul, rN);
    }
    pj += 2 * row;
    for (int k = kbeg; k < kend; k += 2) {
      double p3[3] = {0.0, 0.0, 0.0}, p4[3] = {0.0, 0.0, 0.0};
      if (k + 3 <= kend) load(pj + row, p3);
      link2(k, p1, dO, rO, uO, lO, dN, rN, uN, lN);
      if (k + 4 <= kend) load(pj + 2 * row, p4);
      link2(k + 1, p2, dO, rO, uO, lO, dN, rN, uN, lN);
      pj += 2 * row;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        p1[d] = p3[d];
        p2[d] = p4[d];
      }
    }
  }
}

// Waves per SIMD the register allocation of the lean kernel targets: four blocks of four
// waves per CU.
// The generic kernel (Coulomb fits, log grids, long range) holds more state per lane: at
// two waves per SIMD (256 VGPRs) it runs without spills, and faster than at three or four.
#ifndef PIMC_GENERIC_WAVES
#define PIMC_GENERIC_WAVES 2
#endif
#ifndef PIMC_BISECTION_WAVES
#define PIMC_BISECTION_WAVES 4
#endif
// BisectionStageClass::Sample for midpoint b of the stage (BisectionStage.cc:128-289):
// NEW midpoint = rbar + Delta, Delta ~ N(0, sigma^2) per dimension, and the
// log of the sampling probabilities of the NEW and OLD midpoints.
__device__ inline void sample_midpoint(const PIMC_LDS double* gauss, int b, int skip, double sigma, double sigma2,
                                       double bx0, double bx1, double bx2, double bi0, double bi1, double bi2,
                                       double ip0, double ip1, double ip2, int num_images, PIMC_LDS double* mold,
                                       PIMC_LDS double* mnew, PIMC_LDS double* lnew, PIMC_LDS double* lold) {
  const int s = b * skip, half = skip >> 1;
  // gaussians 3b .. 3b+2 of this (move, level): elements 3b.. of the level's pair sequence
  const double gd[3] = {gauss[3 * b], gauss[3 * b + 1], gauss[3 * b + 2]};
  const double bxs[3] = {bx0, bx1, bx2}, bis[3] = {bi0, bi1, bi2}, ips[3] = {ip0, ip1, ip2};
  double GaussProd = 1.0, GaussProdOld = 1.0, expo = 0.0, expoOld = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double bx = bxs[d], bi = bis[d], ip = ips[d];
    double rdiffOld = put_in_box_1(mold[3 * (s + skip) + d] - mold[3 * s + d], bx, bi, ip);
    double rbarOld = mold[3 * s + d] + 0.5 * rdiffOld;
    double DeltaOld = put_in_box_1(mold[3 * (s + half) + d] - rbarOld, bx, bi, ip);
    double rdiff = put_in_box_1(mnew[3 * (s + skip) + d] - mnew[3 * s + d], bx, bi, ip);
    double rbar = mnew[3 * s + d] + 0.5 * rdiff;
    double Delta = put_in_box_1(gd[d] * sigma, bx, bi, ip);
    if (num_images == 0) {
      // log of a product of single-image Gaussians: the sum of their exponents
      expo += -0.5 * Delta * Delta / sigma2;
      expoOld += -0.5 * DeltaOld * DeltaOld / sigma2;
    } else {
      double GaussSum = 0.0, GaussSumOld = 0.0;
      for (int image = -num_images; image <= num_images; image++) {
        double dist = Delta + (double)image * bx;
        double distOld = DeltaOld + (double)image * bx;
        GaussSum += exp(-0.5 * dist * dist / sigma2);
        GaussSumOld += exp(-0.5 * distOld * distOld / sigma2);
      }
      GaussProd *= GaussSum;
      GaussProdOld *= GaussSumOld;
    }
    mnew[3 * (s + half) + d] = rbar + Delta;
  }
  lnew[b] = num_images == 0 ? expo : log(GaussProd);
  lold[b] = num_images == 0 ? expoOld : log(GaussProdOld);
}

// The Gaussian pairs of every level of a move and the Metropolis draws log(u),
// computed by the whole block at the start of the move (they are pure functions
// of (seed, stream, move, level, index)).  Level l's pairs start at
// gauss_offset(W, l) doubles; level l needs the pairs holding gaussians 0 .. 3 nb - 1.
__host__ __device__ inline int gauss_pairs(int W, int level) {
  const int nb = W >> (level + 1);
  return ((3 * nb - 1) >> 1) + 1;
}
__host__ __device__ inline int gauss_offset(int W, int level) {
  int off = 0;
  for (int l = 0; l < level; l++) off += 2 * gauss_pairs(W, l);
  return off;
}

// KineticClass::SingleAction of one link (KineticClass.cc:33-90).
__device__ inline double kinetic_link_ool(const PIMC_LDS double* a, const PIMC_LDS double* b, double bx0,
                                          double bx1, double bx2, double bi0, double bi1, double bi2,
                                          double ip0, double ip1, double ip2, int num_images,
                                          double FourLambdaTauInv) {
  const double bxs[3] = {bx0, bx1, bx2}, bis[3] = {bi0, bi1, bi2}, ips[3] = {ip0, ip1, ip2};
  double GaussProd = 1.0, expo = 0.0;
#pragma unroll
  for (int dim = 0; dim < 3; dim++) {
    double vel = put_in_box_1(b[dim] - a[dim], bxs[dim], bis[dim], ips[dim]);
    if (num_images == 0) {
      expo += vel * vel * FourLambdaTauInv;   // -log(exp(-x)) summed over dimensions
    } else {
      double GaussSum = 0.0;
      for (int image = -num_images; image <= num_images; image++) {
        double dist = vel + (double)image * bxs[dim];
        GaussSum += exp(-dist * dist * FourLambdaTauInv);
      }
      GaussProd *= GaussSum;
    }
  }
  return num_images == 0 ? expo : -log(GaussProd);
}

// log(u) of the stage's Metropolis draw.
__device__ inline double log_accept_draw(uint64_t seed, uint32_t stream, uint32_t move, int level) {
  return log(rng_uniform(seed, stream, move, kTagAccept, level));
}

// Kernel variants: kLeanDavid -- DavidPA fits on LINEAR grids only, no
// long-range action (the Lennard-Jones systems); kGeneric -- any fits and
// the long-range action, chosen at run time.
enum { kLeanDavid = 0, kGeneric = 1 };

// The moves of replica a.rep_ids[item] (one block).
// Byte offset of the control words (slice1, ptcl, accept, queue item) in the block's LDS.
__host__ __device__ inline size_t bisection_ctl_offset(const LaunchBisection& a) {
  const int W = 1 << a.num_levels;
  return a.lds_table_bytes + sizeof(double) * (6 * (W + 1) + W + 2 * W + 2 * kMaxWaves);
}

// Phase rows (per-dimension e^{i n phi}, both copies) of the slices t = (c0 + j) step,
// j = 0 .. nc-1, into lrC: row pair j holds copy NEW at (2j) stride, OLD at (2j+1)
// stride.  One lane per (slice, copy, dimension).
template <typename LrT>
__device__ inline void phase_rows_chunk(const RepCore& R, const PIMC_LDS double* mnew, const PIMC_LDS double* mold,
                                        int c0, int nc, int step, int stride, PIMC_LDS PhaseRow<LrT>* lrC) {
  for (int i = threadIdx.x; i < nc * 6; i += blockDim.x) {
    const int j = i / 6, c = (i % 6) / 3, d = i % 3;
    const PIMC_LDS double* src = (c == 0 ? mnew : mold) + 3 * (c0 + j) * step;
    const double x[3] = {src[0], src[1], src[2]};
    phase_rows_lane<LrT>(R, x, d, lrC + (2 * j + c) * stride);
  }
}

// PathClass::MoveJoin(oldJoin, newJoin) (PathClass.cc:727-793) by one block: for
// newJoin > oldJoin, OLD(s, p) = NEW(s, Permutation(p)) over s in (oldJoin, newJoin];
// for newJoin < oldJoin, OLD(s, Permutation(p)) = NEW(s, p) over s in (newJoin,
// oldJoin]; then NEW = OLD over those slices.  Outside a move the two copies and
// their permutations are equal.  Called by every thread of the block.
__device__ inline void move_join_block(const RepCore& R, int oldJoin, int newJoin) {
  if (newJoin == oldJoin) return;
  const bool fwd = newJoin > oldJoin;
  const int lo = (fwd ? oldJoin : newJoin) + 1, hi = fwd ? newJoin : oldJoin;
  const int total = (hi - lo + 1) * R.N;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int slice = lo + t / R.N, p = t % R.N;
    const int q = R.perm1[p];
    double v[3];
    load_bead(R, 1, slice, fwd ? q : p, v);
    store_bead(R, 0, slice, fwd ? p : q, v);
  }
  __syncthreads();   // workgroup-scope fence: the OLD writes are visible to the block
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int slice = lo + t / R.N, p = t % R.N;
    double v[3];
    load_bead(R, 0, slice, p, v);
    store_bead(R, 1, slice, p, v);
  }
  __syncthreads();
}

// ---- Several blocks per replica (a cluster) ----
// The `cluster` blocks of one replica run the same moves in lock step: each
// samples the same midpoints (same draws), takes its share of the partners and
// k-vectors, and the members exchange their partial stage sums.  Every member then
// holds the same sums, added in member order, and takes the same Metropolis
// decision.  Every member also writes the accepted beads itself (the same values),
// so no member reads a bead another block wrote.
//
// The exchange needs no cache maintenance: the partial sums are written with
// device-scope atomic stores (write-through to the coherent level), the writer
// waits for those stores to complete before it counts its arrival, and readers
// load the sums with device-scope atomic loads after seeing every arrival.  (An
// agent-scope release/acquire pair would write back and invalidate a whole L2
// at every stage.)

// Thread 0 of every member counts its arrival on the replica's counter and waits
// until all members have arrived n times.  The wait is bounded (a.wait_ticks of
// the device wall clock, or until another block gives up): on a timeout the
// launch's abort flag is set and false is returned, so no wave waits forever.
__device__ inline bool cluster_barrier(const LaunchBisection& a, int item, int n) {
  int* bar = a.bar + item;
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);   // this wave's earlier stores have completed
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int target = a.cluster * n;
  const uint64_t t0 = wall_clock64();
  while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (wall_clock64() - t0 > a.wait_ticks) {
      __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

// Replaces v[0..3] (this member's {SR old, SR new, LR old, LR new}, valid in lane 0) by
// the sums over the members, in member order.  Slots alternate between two buffers: a
// member can only overwrite a buffer after every member has passed the next barrier, that
// is, after every member has read it.  Called by the 64 lanes of wave 0: lane 0 stores and
// waits at the barrier; then lane j loads value j of the 4 x cluster partial sums, so the
// device-coherent loads are in flight together (one round trip instead of one per value),
// and every lane adds them in member order.
__device__ inline bool cluster_sum(const LaunchBisection& a, int item, int member, int nbar, double v[4]) {
  const int lane = (int)threadIdx.x & 63;
  double* slot = a.part + ((size_t)item * 2 + (nbar & 1)) * a.cluster * 4;
  int ok = 1;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++)
      __hip_atomic_store(slot + member * 4 + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok = cluster_barrier(a, item, nbar + 1) ? 1 : 0;
  }
  if (__shfl(ok, 0) == 0) return false;
  const int nv = 4 * a.cluster;   // at most 4 * kMaxCluster = 128 values: two per lane
  const double x0 = lane < nv ? __hip_atomic_load(slot + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
  const double x1 =
      lane + 64 < nv ? __hip_atomic_load(slot + 64 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (int m = 0; m < a.cluster; m++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int j = 4 * m + i;
      t[i] += j < 64 ? __shfl(x0, j) : __shfl(x1, j - 64);
    }
#pragma unroll
  for (int i = 0; i < 4; i++) v[i] = t[i];
  return true;
}

// This member's share [b, e) of the range [lo, hi).
__device__ inline void cluster_share(int lo, int hi, int member, int cluster, int& b, int& e) {
  const int per = (hi - lo + cluster - 1) / cluster;
  b = min(hi, lo + member * per);
  e = min(hi, b + per);
}

template <int NT, bool MIXED, int VAR, bool CL>
__device__ inline void bisection_replica(const LaunchBisection& a, int item, int member, uint32_t first_move,
                                         PIMC_LDS unsigned char* smem) {
  constexpr bool HAS_BC = VAR == kGeneric, HAS_LR = VAR == kGeneric, LIN = VAR == kLeanDavid;
  const ReplicaDev* RD = a.reps + a.rep_ids[item];
  const int pos = a.rep_pos ? a.rep_pos[item] : item;   // the caller's index of this replica
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int L = a.num_levels;
  const int W = 1 << L;
  const int sp = a.species;
  const int tid = threadIdx.x;

  PIMC_LDS unsigned char* tbl = smem;
  PIMC_LDS double* mold = (PIMC_LDS double*)(smem + a.lds_table_bytes);   // [W+1][3]
  PIMC_LDS double* mnew = mold + 3 * (W + 1);                              // [W+1][3]
  PIMC_LDS double* lnew = mnew + 3 * (W + 1);                              // [W/2]
  PIMC_LDS double* lold = lnew + W / 2;                                    // [W/2]
  PIMC_LDS double* kold = lold + W / 2;                                    // [W]
  PIMC_LDS double* knew = kold + W;                                        // [W]
  PIMC_LDS double* scratch = knew + W;                                     // [2 kMaxWaves]
  PIMC_LDS int* ctl = (PIMC_LDS int*)(scratch + 2 * kMaxWaves);           // slice1, ptcl, accept, item
  PIMC_LDS double* gauss = (PIMC_LDS double*)(ctl + 4);                   // [gauss_offset(W, L)]
  PIMC_LDS double* logu_tab = gauss + gauss_offset(W, L);                 // [PIMC_MAX_LEVELS]
  typedef typename std::conditional<MIXED, float, double>::type LrT;   // phase-sum arithmetic
  PIMC_LDS PhaseRow<LrT>* lrC = (PIMC_LDS PhaseRow<LrT>*)(logu_tab + PIMC_MAX_LEVELS);   // [lr_chunk][2][kstride]: NEW, OLD rows

  int64_t st_moves = 0, st_acc = 0, st_stages = 0, st_evals = 0, st_kevals = 0;
  int nbar = 0;   // cluster barriers passed (the same count on every member)
  // this member's k-vectors [kb, ke) of the long-range sums and updates
  int kb = 0, ke = R.nk;
  if constexpr (CL) cluster_share(0, R.nk, member, a.cluster, kb, ke);
  const int numLeft = R.nslices - W;
  const double lambda = S.lambda[sp];
  // PathDataClass::Join, moved to slice2 at the start of every move (BisectionBlock.cc:424).
  // MoveJoin copies beads only when the permutation is not the identity; the blocks of a
  // cluster cannot do that together (the host refuses such a launch; this is the backstop).
  int join = __builtin_amdgcn_readfirstlane(*(const PIMC_GLOBAL int*)R.join);
  int nonid = 0;
  for (int p = tid; p < R.N; p += blockDim.x) nonid |= (R.perm0[p] != p) | (R.perm1[p] != p);
  const bool permuted = __syncthreads_or(nonid) != 0;
  if (CL && permuted) {
    if (tid == 0) __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }

  for (int m = 0; m < a.num_moves; m++) {
    const uint32_t move = first_move + (uint32_t)m;
    if (tid == 0) {
      ctl[0] = (int)floor(rng_uniform(a.seed, R.stream, move, kTagSlice, 0) * numLeft);
      ctl[1] = (int)floor(rng_uniform(a.seed, R.stream, move, kTagPtcl, 0) * S.count[sp]) + S.first[sp];
    }
    {
      // the move's Gaussian pairs (levels lowest .. L-1) and Metropolis draws, one per thread
      const int npairs_tot = (gauss_offset(W, L) - gauss_offset(W, a.lowest_level)) / 2;
      const int nlev = L - a.lowest_level;
      for (int t = tid; t < npairs_tot + nlev; t += blockDim.x) {
        if (t < npairs_tot) {
          int level = a.lowest_level, p = t;
          while (p >= gauss_pairs(W, level)) {
            p -= gauss_pairs(W, level);
            level++;
          }
          double g0, g1;
          gaussian_pair(a.seed, R.stream, move, level, (uint32_t)p, g0, g1);
          PIMC_LDS double* g = gauss + gauss_offset(W, level) + 2 * p;
          g[0] = g0;
          g[1] = g1;
        } else {
          const int level = a.lowest_level + (t - npairs_tot);
          logu_tab[level] = log_accept_draw(a.seed, R.stream, move, level);
        }
      }
    }
    __syncthreads();
    const int slice1 = ctl[0];
    const int ptcl = ctl[1];
    if (permuted) move_join_block(R, join, slice1 + W);
    join = slice1 + W;
    for (int t = tid; t <= W; t += blockDim.x) {
      double v[3];
      load_bead(R, 0, slice1 + t, ptcl, v);
#pragma unroll
      for (int d = 0; d < 3; d++) mold[3 * t + d] = mnew[3 * t + d] = v[d];
    }
    __syncthreads();

    double prev = 0.0;   // thread 0 only
    bool accepted = true, aborted = false;
    int stage = 0;
    for (int level = L - 1; level >= a.lowest_level; level--, stage++) {
      const int skip = 1 << (level + 1);
      const int half = skip >> 1;
      const int nb = W / skip;
      const int K = W >> level;
      const int step = 1 << level;

      // ---- Sample: new midpoints of this level ----
      const double levelTau = 0.5 * R.tau * skip;
      const double sigma2 = (1.0 * lambda * levelTau);
      const double sigma = sqrt(sigma2);
      for (int b = tid; b < nb; b += blockDim.x)
        sample_midpoint(gauss + gauss_offset(W, level), b, skip, sigma, sigma2, R.box[0], R.box[1], R.box[2],
                        R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2], R.num_images,
                        mold, mnew, lnew, lold);
      __syncthreads();

      // ---- Kinetic action of every link, OLD and NEW ----
      const double FourLambdaTauInv = 1.0 / (4.0 * lambda * (R.tau * (1 << level)));
      for (int k = tid; k < K; k += blockDim.x) {
        kold[k] = kinetic_link_ool(mold + 3 * k * step, mold + 3 * (k + 1) * step, R.box[0], R.box[1], R.box[2],
                                   R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2],
                                   R.num_images, FourLambdaTauInv);
        knew[k] = kinetic_link_ool(mnew + 3 * k * step, mnew + 3 * (k + 1) * step, R.box[0], R.box[1], R.box[2],
                                   R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2],
                                   R.num_images, FourLambdaTauInv);
      }

      // ---- Short-range pair action, OLD and NEW (ShortRangeClass.cc:145-171) ----
      StageSums<MIXED> acc;
      for (int s2 = 0; s2 < R.ns; s2++) {
        const FitDev F = load_uniform(&a.fits[S.fit[sp * R.ns + s2]]);
        const UlongView ul = ulong_for(R, F, sp * R.ns + s2, level);
        int j0 = S.first[s2], j1 = S.first[s2] + S.count[s2];
        if constexpr (CL) cluster_share(j0, j1, member, a.cluster, j0, j1);
        if constexpr (HAS_BC) {
          if (F.kind == 1) {
            const BcView Q{qgrid_of(F), level, MIXED ? F.i32_lev[level] : 0};
            sr_partners<NT, MIXED, true, false>(R, F, LevelTable64<true>{}, Q, ul, j0, j1, SkipOne{ptcl}, slice1, K, step, a.link_split, mold,
                                         mnew, acc);
            continue;
          }
        }
        __syncthreads();
        auto T = [&] {
          if constexpr (MIXED)
            return stage_level32<LIN>(F, level + F.taupos, tbl);
          else
            return stage_level64<LIN>(F, level + F.taupos, tbl);
        }();
        __syncthreads();
        sr_partners<NT, MIXED, false, LIN>(R, F, T, BcView{}, UlongView{nullptr, nullptr}, j0, j1, SkipOne{ptcl}, slice1, K, step, a.link_split, mold, mnew,
                                      acc);
      }
      double sr_old, sr_new;   // sr_new: the NEW total (FP64) or the change (mixed) until the test
      // scratch is written next by the long-range sum (after the phase-row barriers) or by the
      // next stage (after the end-of-stage barrier): no trailing barrier
      block_sum2(acc.a, acc.b, scratch, sr_old, sr_new, false);

      // ---- Long-range action at the lowest level (LongRangeClass.cc:273-371) ----
      // Only the moved species' rho_k changes: NEW = OLD + delta over the window, with
      // delta = e^{ik r_new} - e^{ik r_old} of the moved bead.  The action change is formed
      // from delta directly: u |rho + delta|^2 - u |rho|^2 = u (2 Re(rho* delta) + |delta|^2)
      // for the moved species' own term and 2 u Re(rho_s* delta) for each pair (s, sp), so
      // no species' |rho|^2 is formed twice.  The OLD total, a sum over every species pair,
      // is formed only for the move records.  Copy OLD is read; nothing is written until
      // the move is accepted.
      double lr_kold = 0.0, lr_delta = 0.0;
      const bool with_lr = HAS_LR && R.lr_on && level == a.lowest_level;
      if (HAS_LR && with_lr) {
        const bool want_old = a.records != nullptr;
        const int stride = kstride(R), o1 = koff_of(R, 1), o2 = koff_of(R, 2);
        const int nslc = W / step + 1;   // slices t = 0, step, .., W; phase rows staged lr_chunk at a time
        const PIMC_GLOBAL double* ukt = (const PIMC_GLOBAL double*)R.ulong_k + (size_t)level * R.nk;
        const size_t upair = (size_t)R.lr_levels * R.nk;   // distance between species pairs in ulong_k
        const bool own = (R.lr_mask >> (sp * R.ns + sp)) & 1u;
        double part_o = 0.0, part_d = 0.0;
        for (int c0 = 0; c0 < nslc; c0 += a.lr_chunk) {
          const int nc = min(a.lr_chunk, nslc - c0);
          __syncthreads();
          phase_rows_chunk<LrT>(R, mnew, mold, c0, nc, step, stride, lrC);
          __syncthreads();
          for (int k = kb + tid; k < ke; k += blockDim.x) {
           const PIMC_GLOBAL int* kx = (const PIMC_GLOBAL int*)R.kidx + 4 * k;
           const int ki[3] = {kx[0], kx[1], kx[2]};
           // u_k of the pairs that involve the moved species (zero where a pair has no table)
           const double u_own = own ? ukt[(sp * R.ns + sp) * upair + k] : 0.0;
           double u_x[kMaxSpecies];
#pragma unroll
           for (int s = 0; s < kMaxSpecies; s++) {
             const int pr = s < sp ? s * R.ns + sp : sp * R.ns + s;
             u_x[s] = (s < R.ns && s != sp && ((R.lr_mask >> pr) & 1u)) ? ukt[pr * upair + k] : 0.0;
           }
           for (int j = 0; j < nc; j++) {
            const int t = (c0 + j) * step;
            const double factor = (t == 0 || t == W) ? 0.5 : 1.0;
            const PIMC_GLOBAL d2v* rho = (const PIMC_GLOBAL d2v*)R.rho0 + (size_t)(slice1 + t) * R.ns * R.nk;
            LrT nr, ni, orr, oi;
            kterm<LrT>(lrC + 2 * j * stride, o1, o2, ki, nr, ni);
            kterm<LrT>(lrC + (2 * j + 1) * stride, o1, o2, ki, orr, oi);
            const double dx = (double)nr - (double)orr, dy = (double)ni - (double)oi;
            // the species' rho_k this term needs (a species without a pair to sp only for the record)
            d2v ro[kMaxSpecies], rs = d2v{0.0, 0.0};
            double d = 0.0;
#pragma unroll
            for (int s = 0; s < kMaxSpecies; s++) {
              ro[s] = (s < R.ns && (s == sp || want_old || u_x[s] != 0.0)) ? rho[(size_t)s * R.nk + k] : d2v{0.0, 0.0};
              if (s == sp) rs = ro[s];   // selected, not indexed: sp is a run-time value
              else d += u_x[s] * (ro[s].x * dx + ro[s].y * dy);
            }
            part_d += factor * (u_own * (2.0 * (rs.x * dx + rs.y * dy) + (dx * dx + dy * dy)) + 2.0 * d);
            if (want_old) {
#pragma unroll
              for (int s = 0; s < kMaxSpecies; s++)
                if (s < R.ns && ((R.lr_mask >> (s * R.ns + s)) & 1u))
                  part_o += factor * 0.5 * 2.0 * (ro[s].x * ro[s].x + ro[s].y * ro[s].y) * ukt[(s * R.ns + s) * upair + k];
#pragma unroll
              for (int s1 = 0; s1 < kMaxSpecies; s1++)
#pragma unroll
                for (int s2 = s1 + 1; s2 < kMaxSpecies; s2++)
                  if (s2 < R.ns && ((R.lr_mask >> (s1 * R.ns + s2)) & 1u))
                    part_o += factor * 2.0 * (ro[s1].x * ro[s2].x + ro[s1].y * ro[s2].y) * ukt[(s1 * R.ns + s2) * upair + k];
            }
           }
          }
        }
        block_sum2(part_o, part_d, scratch, lr_kold, lr_delta, false);
      }

      // ---- Metropolis test (StageClass.cc:69-72) ----
      bool cl_ok = true;
      if constexpr (CL) {
        if (tid < 64) {   // wave 0; the sums are valid in thread 0
          double v[4] = {sr_old, sr_new, lr_kold, lr_delta};
          cl_ok = cluster_sum(a, item, member, nbar, v);
          sr_old = v[0];
          sr_new = v[1];
          lr_kold = v[2];
          lr_delta = v[3];
        }
        nbar++;
      }
      if (tid == 0 && !cl_ok) ctl[2] = -1;   // a member gave up waiting: abandon the launch
      if (tid == 0 && cl_ok) {
        double logSampleProb = 0.0, logOldSampleProb = 0.0;
        for (int b = 0; b < nb; b++) {
          logOldSampleProb += 0.0 + lold[b];
          logSampleProb += 0.0 + lnew[b];
        }
        const double logSampleRatio = -logSampleProb + logOldSampleProb;
        double kin_o = 0.0, kin_n = 0.0;
        for (int k = 0; k < K; k++) {
          kin_o += kold[k];
          kin_n += knew[k];
        }
        double oldAction = 0.0, newAction = 0.0;
        oldAction += kin_o;
        oldAction += sr_old;
        newAction += kin_n;
        newAction += StageSums<MIXED>::new_total(sr_old, sr_new);
        double currActionChange = newAction - oldAction;
        if (with_lr) {
          // the k-independent terms are the same on both copies and cancel in the change;
          // only the records carry the totals, so the per-slice constants (dependent
          // loads on thread 0's serial path) are summed only when records are kept
          currActionChange += lr_delta;
          if (a.records) {
            double homo = 0.0, background = 0.0, k0 = 0.0;
            for (int t = 0; t <= W; t += step)
              lr_const_terms(R, S, level, (t == 0 || t == W) ? 0.5 : 1.0, homo, background, k0);
            const double cst = R.use_background ? background : k0;
            oldAction += (lr_kold + homo) + cst;
            newAction += ((lr_kold + lr_delta) + homo) + cst;
          }
          st_kevals += 2 * (int64_t)(W / step + 1) * R.nk * R.lr_terms;
        }
        const double logAcceptProb = logSampleRatio - currActionChange + prev;
        const double logu = logu_tab[level];
        const int ok = logAcceptProb >= logu;
        prev = currActionChange;
        ctl[2] = ok;
        st_stages++;
        st_evals += 2 * (int64_t)(R.N - 1) * K;
        if (a.records && member == 0) {
          pimc_move_record* rec = a.records + (size_t)pos * a.num_moves + m;
          rec->old_action[stage] = oldAction;
          rec->new_action[stage] = newAction;
          rec->log_sample_ratio[stage] = logSampleRatio;
          rec->log_u[stage] = logu;
        }
      }
      __syncthreads();
      if (ctl[2] != 1) {
        accepted = false;
        aborted = ctl[2] < 0;
        stage++;
        break;
      }
    }
    if (CL && aborted) return;

    // ---- AcceptCopy: rho_k OLD <- the updated rho_k; PutInBoxFast on the NEW beads, OLD <- NEW ----
    // The chain reads rho_k OLD only and leaves copy NEW behind; the host refreshes NEW from
    // OLD before any call that uses it.
    if (HAS_LR && accepted && R.lr_on) {
      const int step = 1 << a.lowest_level;
      const int stride = kstride(R), o1 = koff_of(R, 1), o2 = koff_of(R, 2);
      const int nslc = W / step + 1;
      for (int c0 = 0; c0 < nslc; c0 += a.lr_chunk) {
        const int nc = min(a.lr_chunk, nslc - c0);
        __syncthreads();
        phase_rows_chunk<LrT>(R, mnew, mold, c0, nc, step, stride, lrC);
        __syncthreads();
        for (int k = kb + tid; k < ke; k += blockDim.x) {
          const int ki[3] = {R.kidx[4 * k], R.kidx[4 * k + 1], R.kidx[4 * k + 2]};
          for (int j = 0; j < nc; j++) {
            const size_t base = ((size_t)(slice1 + (c0 + j) * step) * R.ns + sp) * R.nk;
            LrT nr, ni, orr, oi;
            kterm<LrT>(lrC + 2 * j * stride, o1, o2, ki, nr, ni);
            kterm<LrT>(lrC + (2 * j + 1) * stride, o1, o2, ki, orr, oi);
            d2v v = R.rho0[base + k];
            v.x += (double)nr;
            v.y += (double)ni;
            v.x -= (double)orr;
            v.y -= (double)oi;
            R.rho0[base + k] = v;
          }
        }
      }
    }
    if (accepted) {
      for (int t = tid; t <= W; t += blockDim.x) {
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = put_in_box_fast_1(mnew[3 * t + d], R.box[d]);
        store_bead(R, 0, slice1 + t, ptcl, v);
        store_bead(R, 1, slice1 + t, ptcl, v);
      }
    }
    if (tid == 0 && member == 0) {
      st_moves++;
      st_acc += accepted ? 1 : 0;
      if (a.records) {
        pimc_move_record* rec = a.records + (size_t)pos * a.num_moves + m;
        rec->slice1 = slice1;
        rec->ptcl = ptcl;
        rec->accepted = accepted ? 1 : 0;
        rec->num_stages = stage;
        for (int s = stage; s < PIMC_MAX_LEVELS; s++) {
          rec->old_action[s] = 0.0;
          rec->new_action[s] = 0.0;
          rec->log_sample_ratio[s] = 0.0;
          rec->log_u[s] = 0.0;
        }
        rec->cycle_length = 0;   // no permutation stage
        for (int q = 0; q < kMaxActive; q++) rec->cycle[q] = 0;
        rec->num_cycles = 0;
        rec->forward_t = rec->reverse_t = rec->accept_prob = rec->psi = 0.0;
      }
    }
    __syncthreads();
  }
  if (tid == 0 && member == 0) {
    *R.join = join;
    int64_t* st = a.stats + (size_t)pos * 5;
    st[0] = st_moves;
    st[1] = st_acc;
    st[2] = st_stages;
    st[3] = st_evals;
    st[4] = st_kevals;
    int64_t* ac = a.acc + (size_t)a.rep_ids[item] * 5;   // one block per replica per launch
    ac[0] += st_moves;
    ac[1] += st_acc;
    ac[2] += st_stages;
    ac[3] += st_evals;
    ac[4] += st_kevals;
  }
}

// Launch modes: kOneBlock -- one block per replica; kPersistent -- a persistent
// grid whose blocks take replicas from the work queue in list order (the host
// lists them longest first) until the list is exhausted; kCluster -- a.cluster
// blocks per replica (blocks c*i .. c*i + c - 1 run replica i).  first_move_dev,
// when set, holds the first move number (a captured sweep advances it on the
// device).
enum { kOneBlock = 0, kPersistent = 1, kCluster = 2 };

template <int NT, bool MIXED, int VAR, int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(VAR == kLeanDavid ? PIMC_BISECTION_WAVES : PIMC_GENERIC_WAVES)))
k_bisection(const LaunchBisection a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  PIMC_LDS unsigned char* smem = (PIMC_LDS unsigned char*)smem_raw;
  PIMC_LDS int* take = (PIMC_LDS int*)(smem + bisection_ctl_offset(a)) + 3;   // control word 3
  const uint32_t first_move = a.first_move_dev ? *a.first_move_dev : a.first_move;
  if constexpr (MODE == kOneBlock) {
    bisection_replica<NT, MIXED, VAR, false>(a, blockIdx.x, 0, first_move, smem);
  } else if constexpr (MODE == kCluster) {
    bisection_replica<NT, MIXED, VAR, true>(a, (int)blockIdx.x / a.cluster, (int)blockIdx.x % a.cluster, first_move,
                                            smem);
  } else {
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) *take = atomicAdd(a.queue, 1);
      __syncthreads();
      const int item = *take;
      if (item >= a.nrep) break;   // uniform: every wave of the block leaves here
      bisection_replica<NT, MIXED, VAR, false>(a, item, 0, first_move, smem);
    }
  }
}

inline size_t bisection_lds_bytes(const LaunchBisection& a) {
  const int W = 1 << a.num_levels;
  return bisection_ctl_offset(a) + 4 * sizeof(int) + sizeof(double) * (gauss_offset(W, a.num_levels) + PIMC_MAX_LEVELS) +
         a.lds_lr_bytes;
}

template <bool MIXED, int VAR, int BLOCK, int MODE>
hipError_t launch_bis_p(int nt, const LaunchBisection& a, int grid, hipStream_t s) {
  const size_t lds = bisection_lds_bytes(a);
  switch (nt) {
    case 0: k_bisection<0, MIXED, VAR, BLOCK, MODE><<<grid, BLOCK, lds, s>>>(a); break;
    case 1: k_bisection<1, MIXED, VAR, BLOCK, MODE><<<grid, BLOCK, lds, s>>>(a); break;
    case 2: k_bisection<2, MIXED, VAR, BLOCK, MODE><<<grid, BLOCK, lds, s>>>(a); break;
    default: k_bisection<3, MIXED, VAR, BLOCK, MODE><<<grid, BLOCK, lds, s>>>(a); break;
  }
  return hipGetLastError();
}

template <bool MIXED, int VAR, int BLOCK>
hipError_t launch_bis_b(int nt, const LaunchBisection& a, int grid, hipStream_t s) {
  if (a.cluster > 1) return launch_bis_p<MIXED, VAR, BLOCK, kCluster>(nt, a, grid, s);
  return a.queue ? launch_bis_p<MIXED, VAR, BLOCK, kPersistent>(nt, a, grid, s)
                 : launch_bis_p<MIXED, VAR, BLOCK, kOneBlock>(nt, a, grid, s);
}

// The launch of one (precision, variant) pair; each pair is instantiated in a file of its
// own (pimc_bisection_{lean,gen}{32,64}.hip), so the four compile in parallel.
template <bool MIXED, int VAR>
hipError_t launch_bis_t(int nt, const LaunchBisection& a, int block, int grid, hipStream_t s) {
  return block == 512 ? launch_bis_b<MIXED, VAR, 512>(nt, a, grid, s) : launch_bis_b<MIXED, VAR, 256>(nt, a, grid, s);
}

// Resident blocks per CU of the bisection kernel for one launch shape.
template <bool MIXED, int VAR, int MODE>
hipError_t occupancy_m(int& n, int nt, int block, size_t lds) {
  auto q = [&](auto kern) { return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, block, lds); };
  if (block == 512)
    return nt == 0   ? q(k_bisection<0, MIXED, VAR, 512, MODE>)
           : nt == 1 ? q(k_bisection<1, MIXED, VAR, 512, MODE>)
           : nt == 2 ? q(k_bisection<2, MIXED, VAR, 512, MODE>)
                     : q(k_bisection<3, MIXED, VAR, 512, MODE>);
  return nt == 0   ? q(k_bisection<0, MIXED, VAR, 256, MODE>)
         : nt == 1 ? q(k_bisection<1, MIXED, VAR, 256, MODE>)
         : nt == 2 ? q(k_bisection<2, MIXED, VAR, 256, MODE>)
                   : q(k_bisection<3, MIXED, VAR, 256, MODE>);
}
template <bool MIXED, int VAR>
hipError_t occupancy_t(int& n, int nt, int block, size_t lds, int mode) {
  return mode == kCluster      ? occupancy_m<MIXED, VAR, kCluster>(n, nt, block, lds)
         : mode == kPersistent ? occupancy_m<MIXED, VAR, kPersistent>(n, nt, block, lds)
                               : occupancy_m<MIXED, VAR, kOneBlock>(n, nt, block, lds);
}

// The four instantiations live in their own files.
#define PIMC_BISECTION_VARIANT(M_, V_)                                                                    \
  template hipError_t launch_bis_t<M_, V_>(int, const LaunchBisection&, int, int, hipStream_t);        \
  template hipError_t occupancy_t<M_, V_>(int&, int, int, size_t, int);
#define PIMC_BISECTION_EXTERN(M_, V_)                                                                     \
  extern template hipError_t launch_bis_t<M_, V_>(int, const LaunchBisection&, int, int, hipStream_t); \
  extern template hipError_t occupancy_t<M_, V_>(int&, int, int, size_t, int);

}  // namespace pimc
