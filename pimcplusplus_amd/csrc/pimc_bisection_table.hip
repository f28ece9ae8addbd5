// pimc_bisection_table.hip -- the GPU-resident bisection chain with PermuteType TABLE.
//
// BisectionBlockClass::MakeMove with TablePermuteStageClass first and last in the stage list
// (PIMC++/src/Moves/BisectionBlock.cc:51-64, 126, 410-451), StepsPerBlock = 1:
//   ChooseTimeSlices, MoveJoin(slice2)                        BisectionBlock.cc:363-374, 424
//   InitBlock: Forw->ConstructHTable                          TablePermuteStage.cc:127-140
//   permute stage, first Attempt: ConstructBosonCycleTable,
//     AttemptPermutation (FindEntry, CycleClass::Apply at slice2)
//                                                             TablePermuteStage.cc:204-230
//   bisection stages moving every particle of the cycle       StageClass.cc:49-95
//   permute stage, last Attempt: Rev->UpdateHTable on the NEW copy, CalcReverseProb, the test
//                                                             TablePermuteStage.cc:232-260
//   AcceptCopy / RejectCopy of beads and permutation          PathClass.cc:797-899
// The HTable updates of TablePermuteStageClass::Accept/Reject only serve later steps of the
// same block, and StepsPerBlock = 1 rebuilds the table every move, so they change nothing.
//
// One block (4 waves) per replica at a time; the grid's blocks take the replicas in turn.  The
// moved species' HTable lives in LDS.  The cycle table is built by the block: the cycles that
// begin (i) or (i, j) form one unit each; the units count their kept entries in parallel, one
// thread turns the counts into offsets, the units write their float weights in the
// reference's order, and one thread forms the float running sums in that order (the order of
// the float additions is part of the result).  The drawn entry is the first whose running sum
// exceeds xi Norm, which is what the reference's binary search over the non-decreasing sums
// returns.  A table larger than the block's scratch falls back to walking the loops with one
// thread, twice (once for Norm, once for the search), so its size is unbounded.
#include <algorithm>

#include "pimc_bisection.hpp"

namespace pimc {

// Gaussian pairs of level l with np moved particles (gaussian 3 (pi nb + b) + d of the level).
__host__ __device__ inline int tbl_gauss_pairs(int W, int level, int np) {
  const int nb = W >> (level + 1);
  return ((3 * np * nb - 1) >> 1) + 1;
}
// Offset of level l's pairs, room for kMaxActive particles per level.
__host__ __device__ inline int tbl_gauss_offset(int W, int level) {
  int off = 0;
  for (int l = 0; l < level; l++) off += 2 * tbl_gauss_pairs(W, l, kMaxActive);
  return off;
}

// LDS of the table chain (byte offsets): the staged spline level, the HTable and its 4-cycle
// bounds, the moved
// particles' beads (OLD and NEW, kMaxActive x (W+1) x 3), the per-midpoint log sample
// probabilities and per-link kinetic actions of every moved particle, reduction scratch,
// control words, the move's Gaussian pairs and Metropolis draws, then cycle-table storage up to
// the block's share of the CU's LDS (a.permute_lds_limit).
struct TableLds {
  size_t h, mx, mold, mnew, lnew, lold, kold, knew, scratch, ctl, gauss, logu, cyc, total;
  int cyc_cap;   // cycle-table entries of the LDS area (12 bytes each)
};
// Waves per SIMD the kernel is compiled for (its register budget): with one wave of each block
// on each SIMD, also the blocks a CU holds at once.
constexpr int kTableWaves = 4;
__host__ __device__ inline TableLds table_lds(const LaunchBisection& a) {
  const int W = 1 << a.num_levels;
  auto up = [](size_t x) { return (x + 15) & ~(size_t)15; };
  TableLds t;
  size_t o = up(a.lds_table_bytes);
  t.h = o;
  o += up(a.lds_htable_bytes);
  t.mx = o;
  o += up(a.lds_htable_bytes);
  t.mold = o;
  o += sizeof(double) * kMaxActive * (W + 1) * 3;
  t.mnew = o;
  o += sizeof(double) * kMaxActive * (W + 1) * 3;
  t.lnew = o;
  o += sizeof(double) * kMaxActive * (W / 2);
  t.lold = o;
  o += sizeof(double) * kMaxActive * (W / 2);
  t.kold = o;
  o += sizeof(double) * kMaxActive * W;
  t.knew = o;
  o += sizeof(double) * kMaxActive * W;
  t.scratch = o;
  o += sizeof(double) * 2 * kMaxWaves;
  t.ctl = o;
  o += sizeof(int) * 32;
  t.gauss = o;
  o += sizeof(double) * tbl_gauss_offset(W, a.num_levels);
  t.logu = o;
  o += sizeof(double) * PIMC_MAX_LEVELS;
  // the rest of the block's LDS holds cycle tables (P, C, particles)
  t.cyc = up(o);
  t.cyc_cap = t.cyc < a.permute_lds_limit ? (int)((a.permute_lds_limit - t.cyc) / 12) : 0;
  t.total = t.cyc + (size_t)12 * t.cyc_cap;
  return t;
}

// PermuteTableClass::ConstructHTable entry (PermuteTableClass.cc:122-150) from the beads
// a = r_i(slice1), b = r_i(slice2), c = r_j(slice2): displacements through PutInBox, dot
// products summed over x, y, z in order, exp in FP64 stored as float, 0 below log(epsilon).
__device__ inline float htable_value(const RepCore& R, const double a[3], const double b[3], const double c[3],
                                     double four_lambda_beta_inv, double log_eps) {
  double dii = 0.0, dij = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma clang fp contract(off)
    const double u = put_in_box_1(b[d] - a[d], R.box[d], R.boxinv[d], R.isper[d]);
    const double w = put_in_box_1(c[d] - a[d], R.box[d], R.boxinv[d], R.isper[d]);
    dii += u * u;
    dij += w * w;
  }
  const double to_exp = (-dij + dii) * four_lambda_beta_inv;
  return to_exp < log_eps ? 0.0f : (float)exp(to_exp);
}

// A cycle of the table: its particles (species-relative) and float weight.
struct CycleHit {
  int len;
  int rep[kMaxActive];
  float P;
};

// Gamma[0..3] (the weights of 1- to 4-cycles) and the cut-off of a TABLE stage.  Fields, not an
// array, so that they stay in registers (no index is computed at run time).
struct Gammas {
  double g0, g1, g2, g3, eps;
  __device__ inline double of(int len) const { return len == 1 ? g0 : len == 2 ? g1 : len == 3 ? g2 : g3; }
};

// One walk of ConstructBosonCycleTable (PermuteTableClass.cc:339-449) with no excluded particle
// and no z focus over the n x n table H: every cycle i < j, k, l in the reference's loop order,
// weight Gamma[len-1] times the H products along the cycle, rounded to float, kept above
// epsilon, its running sum C = P + (the last kept C) in float.  Returns Norm (the last C) and
// the number of kept entries.  With target >= 0 the walk stops at the first kept entry whose
// C exceeds target and returns it in *hit (the last kept entry when none does).  The 4-cycle
// loops that the bounds Mx rule out (see unit_entries) are skipped.  One thread.
__device__ float cycle_walk(const PIMC_LDS float* H, const PIMC_LDS float* Mx, int n, const Gammas& G, double target,
                            int* count, CycleHit* hit) {
#pragma clang fp contract(off)
  int num = 0;
  float last = 0.0f;
  int rep[kMaxActive] = {0, 0, 0, 0};
  // keeps an entry; true when the search is over
  auto keep = [&](int len, float P, float C) {
    last = C;
    num++;
    if (target < 0.0) return false;
    hit->len = len;
#pragma unroll
    for (int q = 0; q < kMaxActive; q++) hit->rep[q] = rep[q];
    hit->P = P;
    return (double)C > target;
  };
  for (int i = 0; i < n; i++) {
    rep[0] = i;
    const float P1 = (float)G.g0;
    if (keep(1, P1, num != 0 ? P1 + last : P1)) goto done;
    for (int j = i + 1; j < n; j++) {   // 2 and higher cycles
      rep[1] = j;
      const double hprod = (double)H[i * n + j];
      const float P2 = (float)(G.g1 * hprod * (double)H[j * n + i]);
      if ((double)P2 > G.eps && keep(2, P2, P2 + last)) goto done;
      if (hprod * G.g2 * G.g3 > G.eps) {
        for (int k = i + 1; k < n; k++) {   // 3 and higher cycles
          if (k == j) continue;
          rep[2] = k;
          const double hprod2 = hprod * (double)H[j * n + k];
          const float P3 = (float)(G.g2 * hprod2 * (double)H[k * n + i]);
          if ((double)P3 > G.eps && keep(3, P3, P3 + last)) goto done;
          if (hprod2 * G.g3 > G.eps && G.g3 * hprod2 * (double)Mx[k * n + i] * (1.0 + 1e-6) > G.eps) {
            for (int l = i + 1; l < n; l++) {
              if (l == j || l == k) continue;
              rep[3] = l;
              const double hprod3 = hprod2 * (double)H[k * n + l];
              const float P4 = (float)(G.g3 * hprod3 * (double)H[l * n + i]);
              if ((double)P4 > G.eps && keep(4, P4, P4 + last)) goto done;
            }
          }
        }
      }
    }
  }
done:
  *count = num;
  return last;
}

// Storage of a built cycle table: the kept entries' float weights P, their float running sums C
// and their particles (one byte each, 0xff beyond the cycle's length).  In LDS when the table
// fits there, else in the block's region of global scratch (cap entries).
struct CycleScratch {
  float* P;
  float* C;
  uint32_t* rep;
  int cap;
};
__host__ __device__ inline size_t cycle_region_bytes(int cap) { return (size_t)cap * 12; }
__device__ inline CycleScratch cycle_scratch(const LaunchBisection& a, int block) {
  unsigned char* base = (unsigned char*)a.permute_scratch + (size_t)block * cycle_region_bytes(a.permute_cap);
  CycleScratch S;
  S.P = (float*)base;
  S.C = S.P + a.permute_cap;
  S.rep = (uint32_t*)(S.C + a.permute_cap);
  S.cap = a.permute_cap;
  return S;
}

// The table in units, in the reference's order: per i, the 1-cycle (i), then per j > i the
// 2-cycle (i, j) followed by one unit per k in (i, n) minus {j}: the 3-cycle (i, j, k) and
// the 4-cycles that extend it.  Particle i owns 1 + m^2 units (m = n - i - 1), n + sum of
// m^2 over m < n in all.
struct UnitPos {
  int i, j, k;   // j < 0: the 1-cycle (i); k < 0: the 2-cycle (i, j)
};
__host__ __device__ inline int cycle_units(int n) { return n + (n - 1) * n * (2 * n - 1) / 6; }
__host__ __device__ inline UnitPos unit_at(int u, int n) {
  int i = 0;
  for (; i < n - 1; i++) {
    const int m = n - i - 1;
    if (u < 1 + m * m) break;
    u -= 1 + m * m;
  }
  if (u == 0) return UnitPos{i, -1, -1};
  const int m = n - i - 1, r = u - 1, j = i + 1 + r / m, kk = r % m;
  if (kk == 0) return UnitPos{i, j, -1};
  const int k = i + kk;
  return UnitPos{i, j, k >= j ? k + 1 : k};
}
__host__ __device__ inline void unit_next(UnitPos& p, int n) {
  if (p.j < 0) {   // after the 1-cycle of i: the 2-cycle (i, i+1), or the next i
    if (p.i + 1 < n) {
      p.j = p.i + 1;
    } else {
      p.i++;
    }
    return;
  }
  int k = p.k < 0 ? p.i + 1 : p.k + 1;
  if (k == p.j) k++;
  if (k < n) {
    p.k = k;
    return;
  }
  p.k = -1;
  if (p.j + 1 < n) {
    p.j++;
  } else {
    p.i++;
    p.j = -1;
  }
}

// Bounds that let a unit skip its 4-cycle loop: Mx(k, i) >= max over l > i of H(k,l) H(l,i)
// (the products of two floats are exact in FP64; the maximum is rounded up to float).  Called
// by every thread.
__device__ inline void cycle_bounds(const PIMC_LDS float* H, PIMC_LDS float* Mx, int n) {
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int k = t / n, i = t % n;
    double m = 0.0;
    for (int l = i + 1; l < n; l++) m = fmax(m, (double)H[k * n + l] * (double)H[l * n + i]);
    float f = (float)m;
    if ((double)f < m) f = __uint_as_float(__float_as_uint(f) + 1u);
    Mx[t] = f;
  }
  __syncthreads();
}

// The kept entries of one unit in the reference's order (the loop body of
// ConstructBosonCycleTable, PermuteTableClass.cc:365-440): f(particles, P).  The 4-cycle loop
// is skipped when G3 hprod2 Mx(k, i) (1 + 1e-6) <= eps: every P4 of the loop is at most
// G3 hprod2 H(k,l) H(l,i) times (1 + 6e-8) for the roundings of its two FP64 products and of
// its conversion to float, so none would pass the reference's test P4 > eps.
template <typename Fn>
__device__ inline void unit_entries(const PIMC_LDS float* H, const PIMC_LDS float* Mx, int n, const Gammas& G,
                                    const UnitPos& p, Fn f) {
#pragma clang fp contract(off)
  const int i = p.i, j = p.j;
  if (j < 0) {
    f((uint32_t)i | 0xffffff00u, (float)G.g0);
    return;
  }
  const uint32_t r01 = (uint32_t)i | ((uint32_t)j << 8);
  const double hprod = (double)H[i * n + j];
  if (p.k < 0) {
    const float P2 = (float)(G.g1 * hprod * (double)H[j * n + i]);
    if ((double)P2 > G.eps) f(r01 | 0xffff0000u, P2);
    return;
  }
  if (!(hprod * G.g2 * G.g3 > G.eps)) return;
  const int k = p.k;
  const double hprod2 = hprod * (double)H[j * n + k];
  const float P3 = (float)(G.g2 * hprod2 * (double)H[k * n + i]);
  const uint32_t r012 = r01 | ((uint32_t)k << 16);
  if ((double)P3 > G.eps) f(r012 | 0xff000000u, P3);
  if (!(hprod2 * G.g3 > G.eps)) return;
  if (G.g3 * hprod2 * (double)Mx[k * n + i] * (1.0 + 1e-6) <= G.eps) return;
  for (int l = i + 1; l < n; l++) {
    if (l == j || l == k) continue;
    const double hprod3 = hprod2 * (double)H[k * n + l];
    const float P4 = (float)(G.g3 * hprod3 * (double)H[l * n + i]);
    if ((double)P4 > G.eps) f(r012 | ((uint32_t)l << 24), P4);
  }
}

// Exclusive prefix sum of one int per thread over the block (whole waves); *total gets the
// sum.  Called by every thread.
__device__ inline int block_exclusive_scan(int v, PIMC_LDS int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int t = wsum[w];
    base += w < wave ? t : 0;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// A built table: its entry count and where its entries are (P == nullptr when it exceeds both
// the LDS area and the global region: the caller walks the loops instead).
struct CycleTable {
  float* P;
  float* C;
  uint32_t* rep;
  int E;
};

// The block builds the cycle table of H: each thread takes a contiguous run of units, counts
// their kept entries, a block scan gives each run its offset, and the runs write their entries
// there, into LDS (lcap entries at lP, lC, lrep) or the global region S.  Called by every thread.
__device__ inline CycleTable cycle_table_block(const PIMC_LDS float* H, PIMC_LDS float* Mx, int n, const Gammas& G,
                                               float* lP, float* lC, uint32_t* lrep, int lcap, const CycleScratch& S,
                                               PIMC_LDS int* wsum) {
  cycle_bounds(H, Mx, n);
  const int U = cycle_units(n);
  const int per = (U + (int)blockDim.x - 1) / (int)blockDim.x;
  const int u0 = min(U, (int)threadIdx.x * per), u1 = min(U, u0 + per);
  int c = 0;
  if (u0 < u1) {
    UnitPos p = unit_at(u0, n);
    for (int u = u0; u < u1; u++, unit_next(p, n)) unit_entries(H, Mx, n, G, p, [&](uint32_t, float) { c++; });
  }
  CycleTable t;
  int e = block_exclusive_scan(c, wsum, &t.E);
  if (t.E <= lcap) {
    t.P = lP;
    t.C = lC;
    t.rep = lrep;
  } else if (t.E <= S.cap) {
    t.P = S.P;
    t.C = S.C;
    t.rep = S.rep;
  } else {
    t.P = nullptr;
    t.C = nullptr;
    t.rep = nullptr;
  }
  if (t.P && u0 < u1) {
    UnitPos p = unit_at(u0, n);
    for (int u = u0; u < u1; u++, unit_next(p, n))
      unit_entries(H, Mx, n, G, p, [&](uint32_t r, float P) {
        t.P[e] = P;
        t.rep[e] = r;
        e++;
      });
  }
  __syncthreads();
  return t;
}

// The float running sums of a stored table in order, C[e] = P[e] + C[e-1]; returns Norm, the
// last.  One thread.
__device__ inline float table_sums(const CycleTable& t) {
  float last = 0.0f;
  int e = 0;
  for (; e + 4 <= t.E; e += 4) {
    const float p0 = t.P[e], p1 = t.P[e + 1], p2 = t.P[e + 2], p3 = t.P[e + 3];
    last = p0 + last;
    t.C[e] = last;
    last = p1 + last;
    t.C[e + 1] = last;
    last = p2 + last;
    t.C[e + 2] = last;
    last = p3 + last;
    t.C[e + 3] = last;
  }
  for (; e < t.E; e++) {
    last = t.P[e] + last;
    t.C[e] = last;
  }
  return last;
}
// The drawn entry: the first whose running sum exceeds target, else the last.  The sums are
// non-decreasing (non-negative weights), so a binary search finds it.  One thread.
__device__ inline CycleHit table_find(const CycleTable& t, double target) {
  int lo = 0, hi = t.E;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((double)t.C[mid] > target)
      hi = mid;
    else
      lo = mid + 1;
  }
  const int e = lo < t.E ? lo : t.E - 1;
  CycleHit hit{};
  const uint32_t r = t.rep[e];
  hit.len = 0;
#pragma unroll
  for (int q = 0; q < kMaxActive; q++) {
    const uint32_t b = (r >> (8 * q)) & 0xffu;
    hit.rep[q] = b == 0xffu ? 0 : (int)b;
    hit.len += b == 0xffu ? 0 : 1;
  }
  hit.P = t.P[e];
  return hit;
}

// ShortRange terms of the pairs among the moved particles (ShortRangeClass.cc:138-169 with
// DoPtcl: a pair of moved particles counted once), OLD and NEW, links of this stage.
// Thread t < npair * K takes pair t / K at link t % K; the pairs go through the full
// calcUsqz form from the fit's global tables.
template <int NT>
__device__ inline void moved_pairs(const RepCore& R, const FitDev& F, int level, const BcView& Q, const UlongView& ul,
                                   int np, int K, int step, const PIMC_LDS double* mold, const PIMC_LDS double* mnew,
                                   int rowlen, double& acc_old, double& acc_new) {
  const int npair = np * (np - 1) / 2;
  for (int t = threadIdx.x; t < npair * K; t += blockDim.x) {
    const int pr = t / K, k = t % K;
    int a = 0, b = 1, c = 0;
    for (int x = 0; x < np; x++)
      for (int y = x + 1; y < np; y++, c++)
        if (c == pr) {
          a = x;
          b = y;
        }
    const int s0 = 3 * k * step, s1 = 3 * (k + 1) * step;
    double uo, un;
    {
      const double ma[3] = {mold[a * rowlen + s0], mold[a * rowlen + s0 + 1], mold[a * rowlen + s0 + 2]};
      const double mb[3] = {mold[a * rowlen + s1], mold[a * rowlen + s1 + 1], mold[a * rowlen + s1 + 2]};
      const double pa[3] = {mold[b * rowlen + s0], mold[b * rowlen + s0 + 1], mold[b * rowlen + s0 + 2]};
      const double pb[3] = {mold[b * rowlen + s1], mold[b * rowlen + s1 + 1], mold[b * rowlen + s1 + 2]};
      uo = F.kind == 1 ? pair_link_u<NT>(R, F, Q, ma, mb, pa, pb, ul)
                       : pair_link_u<NT>(R, F, level_table64(F, level + F.taupos), ma, mb, pa, pb, ul);
    }
    {
      const double ma[3] = {mnew[a * rowlen + s0], mnew[a * rowlen + s0 + 1], mnew[a * rowlen + s0 + 2]};
      const double mb[3] = {mnew[a * rowlen + s1], mnew[a * rowlen + s1 + 1], mnew[a * rowlen + s1 + 2]};
      const double pa[3] = {mnew[b * rowlen + s0], mnew[b * rowlen + s0 + 1], mnew[b * rowlen + s0 + 2]};
      const double pb[3] = {mnew[b * rowlen + s1], mnew[b * rowlen + s1 + 1], mnew[b * rowlen + s1 + 2]};
      un = F.kind == 1 ? pair_link_u<NT>(R, F, Q, ma, mb, pa, pb, ul)
                       : pair_link_u<NT>(R, F, level_table64(F, level + F.taupos), ma, mb, pa, pb, ul);
    }
    acc_old += uo;
    acc_new += un;
  }
}

template <int NT, bool MIXED, int VAR>
__device__ void bisection_replica_table(const LaunchBisection& a, int item, PIMC_LDS unsigned char* smem) {
  constexpr bool HAS_BC = VAR == kGeneric, LIN = VAR == kLeanDavid;
  const ReplicaDev* RD = a.reps + a.rep_ids[item];
  const int pos = a.rep_pos ? a.rep_pos[item] : item;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const PermuteDev Pm = a.permute[item];
  const int L = a.num_levels;
  const int W = 1 << L;
  const int sp = a.species;
  const int tid = threadIdx.x;
  const TableLds lay = table_lds(a);
  PIMC_LDS unsigned char* tbl = smem;
  PIMC_LDS float* H = (PIMC_LDS float*)(smem + lay.h);
  PIMC_LDS float* Mx = (PIMC_LDS float*)(smem + lay.mx);
  PIMC_LDS double* mold = (PIMC_LDS double*)(smem + lay.mold);   // [kMaxActive][W+1][3]
  PIMC_LDS double* mnew = (PIMC_LDS double*)(smem + lay.mnew);
  PIMC_LDS double* lnew = (PIMC_LDS double*)(smem + lay.lnew);   // [np][nb]
  PIMC_LDS double* lold = (PIMC_LDS double*)(smem + lay.lold);
  PIMC_LDS double* kold = (PIMC_LDS double*)(smem + lay.kold);   // [np][K]
  PIMC_LDS double* knew = (PIMC_LDS double*)(smem + lay.knew);
  PIMC_LDS double* scratch = (PIMC_LDS double*)(smem + lay.scratch);
  PIMC_LDS int* ctl = (PIMC_LDS int*)(smem + lay.ctl);
  PIMC_LDS double* gauss = (PIMC_LDS double*)(smem + lay.gauss);
  PIMC_LDS double* logu_tab = (PIMC_LDS double*)(smem + lay.logu);
  PIMC_LDS int* cp = ctl + 2;                              // the cycle's particles (absolute)
  PIMC_LDS int* wsum = ctl + 8;                            // scan scratch, one word per wave
  float* lP = (float*)(PIMC_LDS float*)(smem + lay.cyc);    // LDS cycle-table storage
  float* lC = lP + lay.cyc_cap;
  uint32_t* lrep = (uint32_t*)(lC + lay.cyc_cap);
  PIMC_LDS double* dctl = (PIMC_LDS double*)(ctl + 16);   // forwT, P, prevActionChange
  const int rowlen = 3 * (W + 1);
  const CycleScratch CS = cycle_scratch(a, blockIdx.x);

  int64_t st_moves = 0, st_acc = 0, st_stages = 0, st_evals = 0;
  const int first = S.first[sp], n = S.count[sp];
  const int numLeft = R.nslices - W;
  const double lambda = S.lambda[sp];
  const double beta = R.tau * (double)W;
  const double fourLambdaBetaInv = 1.0 / (4.0 * lambda * beta);
  const double logEps = log(Pm.epsilon);
  const Gammas Gm{Pm.gamma[0], Pm.gamma[1], Pm.gamma[2], Pm.gamma[3], Pm.epsilon};
  int join = __builtin_amdgcn_readfirstlane(*(const PIMC_GLOBAL int*)R.join);
  int nonid = 0;
  for (int p = tid; p < R.N; p += blockDim.x) nonid |= (R.perm0[p] != p) | (R.perm1[p] != p);
  bool permuted = __syncthreads_or(nonid) != 0;

  for (int m = 0; m < a.num_moves; m++) {
    const uint32_t move = a.first_move + (uint32_t)m;
    if (tid == 0) ctl[0] = (int)floor(rng_uniform(a.seed, R.stream, move, kTagSlice, 0) * numLeft);
    __syncthreads();
    const int slice1 = ctl[0], slice2 = slice1 + W;
    if (permuted) move_join_block(R, join, slice2);
    join = slice2;
    // InitBlock: the forward HTable of the window (both copies are equal here)
    for (int t = tid; t < n * n; t += blockDim.x) {
      const int i = t / n, j = t % n;
      double ai[3], bi[3], cj[3];
      load_bead(R, 0, slice1, first + i, ai);
      load_bead(R, 0, slice2, first + i, bi);
      load_bead(R, 0, slice2, first + j, cj);
      H[t] = htable_value(R, ai, bi, cj, fourLambdaBetaInv, logEps);
    }
    __syncthreads();
    // first Attempt of the permute stage: the cycle table, the drawn cycle
    const CycleTable ft = cycle_table_block(H, Mx, n, Gm, lP, lC, lrep, lay.cyc_cap, CS, wsum);
    if (tid == 0) {
      int count = ft.E;
      CycleHit hit{};
      const double xi = rng_uniform(a.seed, R.stream, move, kTagPermute, 0);
      double Norm;
      if (ft.P) {
        Norm = table_sums(ft);
        hit = table_find(ft, xi * Norm);
      } else {   // beyond the storage: walk the loops twice
        Norm = cycle_walk(H, Mx, n, Gm, -1.0, &count, &hit);
        int count2 = 0;
        cycle_walk(H, Mx, n, Gm, xi * Norm, &count2, &hit);
      }
      ctl[1] = hit.len;
      for (int q = 0; q < kMaxActive; q++) ctl[2 + q] = hit.rep[q] + first;
      ctl[6] = count;
      const double forwT = hit.P * (1.0 / Norm);
      dctl[0] = forwT;
      dctl[1] = (double)hit.P;
      dctl[2] = -log(hit.P / Gm.of(hit.len));   // prevActionChange
    }
    {
      // the move's Gaussian pairs (np particles) and Metropolis draws
      __syncthreads();
      const int np = ctl[1];
      const int nlev = L - a.lowest_level;
      int npairs_tot = 0;
      for (int l = a.lowest_level; l < L; l++) npairs_tot += tbl_gauss_pairs(W, l, np);
      for (int t = tid; t < npairs_tot + nlev; t += blockDim.x) {
        if (t < npairs_tot) {
          int level = a.lowest_level, p = t;
          while (p >= tbl_gauss_pairs(W, level, np)) {
            p -= tbl_gauss_pairs(W, level, np);
            level++;
          }
          double g0, g1;
          gaussian_pair(a.seed, R.stream, move, level, (uint32_t)p, g0, g1);
          PIMC_LDS double* g = gauss + tbl_gauss_offset(W, level) + 2 * p;
          g[0] = g0;
          g[1] = g1;
        } else {
          const int level = a.lowest_level + (t - npairs_tot);
          logu_tab[level] = log(rng_uniform(a.seed, R.stream, move, kTagAccept, level));
        }
      }
    }
    const int np = ctl[1];
    const SkipSet cyc{{cp[0], cp[1], cp[2], cp[3]}, np};   // indexed with constants only
    // the moved particles' window beads; CycleClass::Apply at slice2 on the NEW copy:
    // particle rep[i] takes the bead of rep[i+1], the last that of rep[0]
    for (int t = tid; t < np * (W + 1); t += blockDim.x) {
      const int pi = t / (W + 1), s = t % (W + 1);
      double v[3], w[3];
      load_bead(R, 0, slice1 + s, cp[pi], v);
      if (s == W) load_bead(R, 0, slice2, cp[(pi + 1) % np], w);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        mold[pi * rowlen + 3 * s + d] = v[d];
        mnew[pi * rowlen + 3 * s + d] = s == W ? w[d] : v[d];
      }
    }
    __syncthreads();

    double prev = dctl[2];   // thread 0 only
    bool accepted = true;
    int stage = 0;
    for (int level = L - 1; level >= a.lowest_level; level--, stage++) {
      const int skip = 1 << (level + 1);
      const int nb = W / skip;
      const int K = W >> level;
      const int step = 1 << level;
      // ---- Sample: new midpoints of every moved particle ----
      const double levelTau = 0.5 * R.tau * skip;
      const double sigma2 = (1.0 * lambda * levelTau);
      const double sigma = sqrt(sigma2);
      for (int t = tid; t < np * nb; t += blockDim.x) {
        const int pi = t / nb, b = t % nb;
        sample_midpoint(gauss + tbl_gauss_offset(W, level) + 3 * pi * nb, b, skip, sigma, sigma2, R.box[0], R.box[1],
                        R.box[2], R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2],
                        R.num_images, mold + pi * rowlen, mnew + pi * rowlen, lnew + pi * nb, lold + pi * nb);
      }
      __syncthreads();
      // ---- Kinetic action of every link of every moved particle ----
      const double FourLambdaTauInv = 1.0 / (4.0 * lambda * (R.tau * (1 << level)));
      for (int t = tid; t < np * K; t += blockDim.x) {
        const int pi = t / K, k = t % K;
        const PIMC_LDS double* mo = mold + pi * rowlen;
        const PIMC_LDS double* mn = mnew + pi * rowlen;
        kold[t] = kinetic_link_ool(mo + 3 * k * step, mo + 3 * (k + 1) * step, R.box[0], R.box[1], R.box[2],
                                   R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2],
                                   R.num_images, FourLambdaTauInv);
        knew[t] = kinetic_link_ool(mn + 3 * k * step, mn + 3 * (k + 1) * step, R.box[0], R.box[1], R.box[2],
                                   R.boxinv[0], R.boxinv[1], R.boxinv[2], R.isper[0], R.isper[1], R.isper[2],
                                   R.num_images, FourLambdaTauInv);
      }
      // ---- Short range, OLD and NEW: every moved particle against the unmoved ones, then the
      // pairs among the moved ones ----
      StageSums<MIXED> acc;
      double mp_old = 0.0, mp_new = 0.0;
      for (int s2 = 0; s2 < R.ns; s2++) {
        const FitDev F = load_uniform(&a.fits[S.fit[sp * R.ns + s2]]);
        const UlongView ul = ulong_for(R, F, sp * R.ns + s2, level);
        const int j0 = S.first[s2], j1 = S.first[s2] + S.count[s2];
        if (s2 == sp) moved_pairs<NT>(R, F, level, BcView{qgrid_of(F), level, 0}, ul, np, K, step, mold, mnew, rowlen,
                                      mp_old, mp_new);
        if constexpr (HAS_BC) {
          if (F.kind == 1) {
            const BcView Q{qgrid_of(F), level, MIXED ? F.i32_lev[level] : 0};
            for (int pi = 0; pi < np; pi++)
              sr_partners<NT, MIXED, true, false>(R, F, LevelTable64<true>{}, Q, ul, j0, j1, cyc, slice1, K, step,
                                                  a.link_split, mold + pi * rowlen, mnew + pi * rowlen, acc);
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
        for (int pi = 0; pi < np; pi++)
          sr_partners<NT, MIXED, false, LIN>(R, F, T, BcView{}, UlongView{nullptr, nullptr}, j0, j1, cyc, slice1, K,
                                             step, a.link_split, mold + pi * rowlen, mnew + pi * rowlen, acc);
      }
      double sr_old, sr_new;   // sr_new: the NEW total (FP64) or the change (mixed)
      block_sum2(acc.a + mp_old, acc.b + StageSums<MIXED>::b_of(mp_old, mp_new), scratch, sr_old, sr_new);

      // ---- Metropolis test (StageClass.cc:69-72) ----
      if (tid == 0) {
        double logSampleProb = 0.0, logOldSampleProb = 0.0;
        for (int b = 0; b < np * nb; b++) {
          logOldSampleProb += 0.0 + lold[b];
          logSampleProb += 0.0 + lnew[b];
        }
        const double logSampleRatio = -logSampleProb + logOldSampleProb;
        double kin_o = 0.0, kin_n = 0.0;
        for (int k = 0; k < np * K; k++) {
          kin_o += kold[k];
          kin_n += knew[k];
        }
        double oldAction = 0.0, newAction = 0.0;
        oldAction += kin_o;
        oldAction += sr_old;
        newAction += kin_n;
        newAction += StageSums<MIXED>::new_total(sr_old, sr_new);
        const double currActionChange = newAction - oldAction;
        const double logAcceptProb = logSampleRatio - currActionChange + prev;
        const double logu = logu_tab[level];
        const int ok = logAcceptProb >= logu;
        prev = currActionChange;
        ctl[7] = ok;
        st_stages++;
        st_evals += 2 * (int64_t)(np * (R.N - np) + np * (np - 1) / 2) * K;
        if (a.records) {
          pimc_move_record* rec = a.records + (size_t)pos * a.num_moves + m;
          rec->old_action[stage] = oldAction;
          rec->new_action[stage] = newAction;
          rec->log_sample_ratio[stage] = logSampleRatio;
          rec->log_u[stage] = logu;
        }
      }
      __syncthreads();
      if (ctl[7] != 1) {
        accepted = false;
        stage++;
        break;
      }
    }

    // ---- last Attempt of the permute stage: the reverse table and the test ----
    double revT = 0.0, acceptProb = 0.0, psi = 0.0;
    if (accepted) {
      if (np != 1) {
        // Rev = Forw's table with the rows and columns of the cycle's particles recomputed on
        // the NEW copy (UpdateHTable): slice1 beads are unchanged, the slice2 beads of the
        // cycle are the applied ones
        for (int t = tid; t < n * n; t += blockDim.x) {
          const int i = t / n, j = t % n;
          int ci = -1, cj = -1;
          for (int q = 0; q < np; q++) {
            ci = cp[q] == first + i ? q : ci;
            cj = cp[q] == first + j ? q : cj;
          }
          if (ci < 0 && cj < 0) continue;
          double ai[3], bi[3], c[3];
          load_bead(R, 0, slice1, first + i, ai);
          load_bead(R, 0, slice2, first + i, bi);
          load_bead(R, 0, slice2, first + j, c);
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if (ci >= 0) bi[d] = mnew[ci * rowlen + 3 * W + d];
            if (cj >= 0) c[d] = mnew[cj * rowlen + 3 * W + d];
          }
          H[t] = htable_value(R, ai, bi, c, fourLambdaBetaInv, logEps);
        }
        __syncthreads();
      }
      CycleTable rt{nullptr, nullptr, nullptr, 0};
      if (np != 1) rt = cycle_table_block(H, Mx, n, Gm, lP, lC, lrep, lay.cyc_cap, CS, wsum);
      if (tid == 0) {
        const double P = dctl[1];
        const double forwT = dctl[0];
        if (np != 1) {
          int count = 0;
          CycleHit unused{};
          const double revNorm = rt.P ? (double)table_sums(rt)
                                      : (double)cycle_walk(H, Mx, n, Gm, -1.0, &count, &unused);
          revT = (Gm.of(np) * Gm.of(np)) / (P) * (1.0 / revNorm);
        } else {
          revT = forwT;
        }
        const double actionChange = -log(P / Gm.of(np));
        const double pi_ratio = exp(-actionChange);
        const double Tratio = forwT / revT;
        acceptProb = fmin(1.0, pi_ratio / Tratio);
        psi = rng_uniform(a.seed, R.stream, move, kTagPermuteAccept, 0);
        ctl[7] = acceptProb > psi;
      }
      __syncthreads();
      accepted = ctl[7] == 1;
    }

    // ---- AcceptCopy: PutInBoxFast on the NEW beads, OLD <- NEW, permutation OLD <- NEW ----
    if (accepted) {
      for (int t = tid; t < np * (W + 1); t += blockDim.x) {
        const int pi = t / (W + 1), s = t % (W + 1);
        double v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = put_in_box_fast_1(mnew[pi * rowlen + 3 * s + d], R.box[d]);
        store_bead(R, 0, slice1 + s, cp[pi], v);
        store_bead(R, 1, slice1 + s, cp[pi], v);
      }
      if (tid < np && np > 1) {   // Permutation(rep[i]) <- Permutation(rep[i+1]), both copies
        const int v = R.perm1[cp[(tid + 1) % np]];
        __builtin_amdgcn_wave_barrier();
        R.perm0[cp[tid]] = v;
        R.perm1[cp[tid]] = v;
      }
      permuted = permuted || np > 1;
    }
    if (tid == 0) {
      st_moves++;
      st_acc += accepted ? 1 : 0;
      if (a.records) {
        pimc_move_record* rec = a.records + (size_t)pos * a.num_moves + m;
        rec->slice1 = slice1;
        rec->ptcl = cp[0];
        rec->accepted = accepted ? 1 : 0;
        rec->num_stages = stage;
        for (int s = stage; s < PIMC_MAX_LEVELS; s++) {
          rec->old_action[s] = 0.0;
          rec->new_action[s] = 0.0;
          rec->log_sample_ratio[s] = 0.0;
          rec->log_u[s] = 0.0;
        }
        rec->cycle_length = np;
        for (int q = 0; q < kMaxActive; q++) rec->cycle[q] = q < np ? cp[q] : 0;
        rec->num_cycles = ctl[6];
        rec->forward_t = dctl[0];
        rec->reverse_t = revT;
        rec->accept_prob = acceptProb;
        rec->psi = psi;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    *R.join = join;
    int64_t* st = a.stats + (size_t)pos * 5;
    st[0] = st_moves;
    st[1] = st_acc;
    st[2] = st_stages;
    st[3] = st_evals;
    st[4] = 0;
    int64_t* ac = a.acc + (size_t)a.rep_ids[item] * 5;
    ac[0] += st_moves;
    ac[1] += st_acc;
    ac[2] += st_stages;
    ac[3] += st_evals;
  }
}

template <int NT, bool MIXED, int VAR>
__global__ void __launch_bounds__(kBlock, kTableWaves) k_bisection_table(const LaunchBisection a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  for (int item = blockIdx.x; item < a.nrep; item += gridDim.x) {   // uniform per block
    bisection_replica_table<NT, MIXED, VAR>(a, item, (PIMC_LDS unsigned char*)smem_raw);
    __syncthreads();
  }
}

size_t bisection_table_scratch_bytes(int cap, int grid) { return cycle_region_bytes(cap) * grid; }

size_t bisection_table_lds_bytes(const LaunchBisection& a) { return table_lds(a).total; }

size_t bisection_table_lds_limit(size_t lds_per_cu) {
  // 1 KB of each share is left for the kernel's static LDS (reduction words)
  const size_t share = lds_per_cu / kTableWaves;
  return std::min<size_t>(64 * 1024, (share > 1024 ? share - 1024 : 0) & ~(size_t)255);
}

template <bool MIXED, int VAR>
static hipError_t table_occupancy_t(int& n, int nt, size_t lds) {
  switch (nt) {
    case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_bisection_table<0, MIXED, VAR>, kBlock, lds);
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_bisection_table<1, MIXED, VAR>, kBlock, lds);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_bisection_table<2, MIXED, VAR>, kBlock, lds);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_bisection_table<3, MIXED, VAR>, kBlock, lds);
  }
}

int bisection_table_blocks_per_cu(int num_terms, bool mixed, bool lean, size_t lds) {
  int n = 0;
  const int nt = num_terms > 3 ? 3 : num_terms;
  const hipError_t e = lean ? (mixed ? table_occupancy_t<true, kLeanDavid>(n, nt, lds)
                                     : table_occupancy_t<false, kLeanDavid>(n, nt, lds))
                            : (mixed ? table_occupancy_t<true, kGeneric>(n, nt, lds)
                                     : table_occupancy_t<false, kGeneric>(n, nt, lds));
  return e == hipSuccess ? n : 0;
}

template <bool MIXED, int VAR>
static hipError_t launch_table_t(int nt, const LaunchBisection& a, size_t lds, hipStream_t s) {
  const int grid = a.permute_grid;
  switch (nt) {
    case 0: k_bisection_table<0, MIXED, VAR><<<grid, kBlock, lds, s>>>(a); break;
    case 1: k_bisection_table<1, MIXED, VAR><<<grid, kBlock, lds, s>>>(a); break;
    case 2: k_bisection_table<2, MIXED, VAR><<<grid, kBlock, lds, s>>>(a); break;
    default: k_bisection_table<3, MIXED, VAR><<<grid, kBlock, lds, s>>>(a); break;
  }
  return hipGetLastError();
}

hipError_t launch_bisection_table(int num_terms, bool mixed, bool lean, const LaunchBisection& a, hipStream_t s) {
  if (a.nrep <= 0 || a.num_moves <= 0) return hipSuccess;
  const size_t lds = table_lds(a).total;
  if (lds > 64 * 1024 || !a.permute || a.num_levels > 5) return hipErrorInvalidValue;
  if (!a.permute_scratch || a.permute_grid < 1 || a.permute_grid > a.nrep || a.permute_cap < 1) return hipErrorInvalidValue;
  if (a.link_split != 1 && a.link_split != 2 && a.link_split != 4) return hipErrorInvalidValue;
  const int nt = num_terms > 3 ? 3 : num_terms;
  if (lean) return mixed ? launch_table_t<true, kLeanDavid>(nt, a, lds, s) : launch_table_t<false, kLeanDavid>(nt, a, lds, s);
  return mixed ? launch_table_t<true, kGeneric>(nt, a, lds, s) : launch_table_t<false, kGeneric>(nt, a, lds, s);
}

}  // namespace pimc
