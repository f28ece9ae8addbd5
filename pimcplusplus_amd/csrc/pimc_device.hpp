// pimc_device.hpp -- device-side building blocks of the action kernels (gfx950).
//
// Pair geometry, the spline-fitted pair action, the kinetic link action and the
// counter-based random stream.  Geometry that decides an integer (periodic
// image, grid interval) is written with floating-point contraction off, so the
// integers are exactly the ones PIMC++'s source computes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pimc {

constexpr int kMaxSpecies = 4;
constexpr int kMaxLevels = 8;
constexpr int kMaxActive = 4;

// Plain vector types for table data (they load from LDS as single ds_read_b128s).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));

#define PIMC_LDS __attribute__((address_space(3)))
#define PIMC_GLOBAL __attribute__((address_space(1)))

// Pointer to const T, in LDS (InLds) or in global memory.
template <bool InLds>
struct Ptr {
  template <typename T>
  using c = const T*;
};
template <>
struct Ptr<true> {
  template <typename T>
  using c = const PIMC_LDS T*;
};

// --------------------------------------------------------------------------
// Device descriptors
// --------------------------------------------------------------------------

// A 1-D grid as the spline lookups see it (LinearGrid / the .dm reader's LogGrid).
struct GridDev {
  int type;           // 0 linear, 1 log
  int n;
  double start, end;
  double deltainv;    // linear
  double r0, logspacing;
  const double* x;    // [n] points
  const d4v* iv;      // [n-1] interval records {X_i, h_i, 1/h_i, X_{i+1}}
  double r0inv;       // log: 1/r0
  float log2inv;      // log: 1/log2(spacing)
};

// One DavidPAClass fit (kind 0) or PAcoulombBCFitClass fit (kind 1).  Per tau level and grid interval i the spline columns
// used by U are stored as cubic polynomials in the local coordinate
// b = (x - X_i) * hinv_i:  value = c0 + b(c1 + b(c2 + b c3)).  Column 1 (u00)
// comes first, then the off-diagonal columns 2..num_cols-1.
struct FitDev {
  int grid_type;      // 0 linear, 1 log
  int ng;             // grid points
  int nint;           // intervals = ng - 1
  int num_tau;
  int ncol;           // stored columns: 1 + noff
  int noff;           // off-diagonal columns used by num_terms
  int num_terms;
  int taupos;
  double start, end;
  double deltainv;    // linear grid
  double r0, spacing; // log grid
  double logspacing;  // log(spacing), computed on the host
  const double* x;        // [ng] grid points
  const d2v* geo;         // [nint] {X_i, 1/(X_{i+1}-X_i)}
  const d4v* c64;         // [num_tau][nint][ncol]
  // [num_tau][nint][ncol] FP32 copy of c64; an interval whose coefficients are too
  // large for FP32 (sum over |c0..c3| of a column above kMixedLimit) holds NaN in
  // every column, and the mixed path evaluates it from c64.
  const f4v* c32;
  // kind 1: bicubic U(q, t) per level; the q grid is the fields above
  int kind;
  int nt;
  GridDev tgrid;
  GridDev qgrid;          // kind 1: the q grid with its interval records
  const d4v* bc;          // [num_tau][ng][nt] BCpoint {z, dzdx, dzdy, d2zdxdy}
  // beta derivatives for the energy estimator (nullptr when not given):
  // kind 0: [num_tau][nint][ecol] cubics of V (column 0 of the table), du00 and the
  // off-diagonal dukj columns; kind 1: dU bicubic nodes like bc
  const d4v* e64;
  int ecol;
  const d4v* dbc;
  double smallest_beta;
  double z1z2;
  // kind 0: the FP32 rows of intervals i >= i32 are valid (not NaN) at every
  // level; the mixed loops evaluate the FP64 rows of intervals below i32_lev[level]
  // kind 1: c32 holds each node's {dz/dx, dz/dy, d2z/dxdy, 0} in FP32 ([num_tau][ng][nt]
  // f4v), and the mixed path evaluates the cells of q intervals below i32_lev[level] from
  // the FP64 nodes
  int i32;
  const int* i32_lev;     // [num_tau] the same bound for each tau level alone
  // intervals whose FP64 rows a mixed stage copies into LDS beside the FP32 rows:
  // i32 when they take at most kFp64StageBytes, else 0 (read from global memory)
  int i32_lds;
};

// LDS a mixed stage may spend on the FP64 rows of the intervals below i32.
#ifndef PIMC_FP64_STAGE_BYTES
#define PIMC_FP64_STAGE_BYTES 16384
#endif
constexpr int kFp64StageBytes = PIMC_FP64_STAGE_BYTES;

// Largest |c0|+|c1|+|c2|+|c3| of a column the mixed path evaluates in FP32, at each tau
// level.  Over 26,000 stages of the benchmark system the largest |dS error| is 3.0e-7 at
// 0.35 (5.3e-7 at 0.5, 2.3e-6 at 1.0).
#ifndef PIMC_MIXED_LIMIT
#define PIMC_MIXED_LIMIT 0.35
#endif
constexpr double kMixedLimit = PIMC_MIXED_LIMIT;

// Largest variation across a Coulomb bicubic cell (sum over its corners of |z - z_first| +
// h|dz/dx| + k|dz/dy| + hk|d2z/dxdy|) the mixed path evaluates in FP32, anchored on the
// cell's first FP64 node value.
#ifndef PIMC_BC_MIXED_LIMIT
#define PIMC_BC_MIXED_LIMIT 4.0
#endif
constexpr double kBcMixedLimit = PIMC_BC_MIXED_LIMIT;

// A knot of a quintic Ulong table for the mixed path: Y in FP64, then B, C, D, E, F in FP32.
struct __attribute__((aligned(32))) UlongKnot32 {
  double y;
  float c[5];
  float pad;
};

// Scalars of a replica the inner loops use; kernels hold a block-uniform copy
// (scalar registers).  Arrays are only indexed with compile-time constants.
struct RepCore {
  int N;              // particles
  int nslices;        // M + 1
  int ns;             // species
  int num_images;
  double box[3], boxinv[3], isper[3];
  double tau;
  uint32_t stream;
  int nk;
  double* path0;      // OLD copy, [slice][3][N] (slice-major, one row per dimension)
  double* path1;      // NEW copy
  int* perm0;         // [N]
  int* perm1;
  int* join;          // the join slice (PathDataClass::Join), one int
  // k space (PathClass::SetupkVecs3D): C_d(n) tables run over n in [-maxk_d, maxk_d]
  int maxk0, maxk1, maxk2;
  int lr_on;          // long-range action evaluated (UseLongRange and tables present)
  double kbox[3];     // 2 pi / box
  const int* kidx;    // [nk][4]: index + maxk per dimension, pad
  d2v* rho0;          // OLD rho_k, [nslices][ns][nk] (re, im)
  d2v* rho1;          // NEW rho_k
  int lr_levels;      // levels stored in the long-range tables
  int use_background;
  const double* ulong_k;   // [ns*ns][lr_levels][nk] (zero rows where a pair has no table)
  const double* lr_const;  // [ns*ns][3][lr_levels]: Ulong_r0, Ushort_k0, Ulong_k0
  unsigned lr_mask;        // bit sa*ns+sb set where pair (sa, sb) has tables
  int lr_terms;            // homologous species + heterologous pairs with tables
  // Ulong(level)(r) quintic splines (one LinearGrid for every pair) for long-ranged fits
  int use_long_range;      // ActionsClass UseLongRange
  int ulong_n, ulong_levels;
  unsigned ulong_mask;     // bit sa*ns+sb set where the pair has Ulong tables
  double ulong_start, ulong_delta, ulong_deltainv;
  const double* ulong;     // [ns*ns][ulong_levels][ulong_n][6]
  const struct UlongKnot32* ulong32;   // the same knots for the mixed path, [ns*ns][ulong_levels][ulong_n]
  const double* dulong;    // dUlong, same layout, or nullptr
  const double* dulong_k;  // [ns*ns][lr_levels][nk] dUlong_k, or nullptr
  const double* dlr_const; // [ns*ns][3][lr_levels]: dUlong_r0, dUshort_k0, dUlong_k0
  // long-range potential (potential-energy estimator), on the Ulong grid: Vlong(r)
  // [ns*ns][ulong_n][6], Vlong_k [ns*ns][nk], {Vlong_r0, Vshort_k0, Vlong_k0} [ns*ns][3]
  const double* vlong;
  const double* vlong_k;
  const double* vlr_const;
  unsigned vlong_mask;     // bit sa*ns+sb set where the pair has Vlong tables
  __device__ __host__ inline double* path(int mode) const { return mode ? path1 : path0; }
  __device__ __host__ inline int* perm(int mode) const { return mode ? perm1 : perm0; }
  __device__ __host__ inline d2v* rho(int mode) const { return mode ? rho1 : rho0; }
};

// Per-species data, read through the descriptor pointer.
struct RepSpecies {
  int first[kMaxSpecies], count[kMaxSpecies];
  double lambda[kMaxSpecies];
  int fit[kMaxSpecies * kMaxSpecies];
};

struct ReplicaDev {
  RepCore core;
  RepSpecies sp;
};

// --------------------------------------------------------------------------
// Philox4x32-10 and the random-stream layout
//   key = seed; counter = {move, tag<<24 | level<<16 | attempt, index, stream}
// --------------------------------------------------------------------------
enum { kTagSlice = 1, kTagPtcl = 2, kTagGauss = 3, kTagAccept = 4, kTagPermute = 5, kTagPermuteAccept = 6 };

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__host__ __device__ inline double uniform53(uint32_t w0, uint32_t w1) {
  uint64_t v = ((uint64_t)w1 << 32) | (uint64_t)w0;
  return (double)(v >> 11) * 0x1.0p-53;
}

__host__ __device__ inline void rng_block(uint64_t seed, uint32_t stream, uint32_t move, int tag, int level,
                                          uint32_t attempt, uint32_t index, uint32_t out[4]) {
  out[0] = move;
  out[1] = ((uint32_t)tag << 24) | (((uint32_t)level & 0xffu) << 16) | (attempt & 0xffffu);
  out[2] = index;
  out[3] = stream;
  philox4x32_10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__host__ __device__ inline double rng_uniform(uint64_t seed, uint32_t stream, uint32_t move, int tag,
                                              int level) {
  uint32_t w[4];
  rng_block(seed, stream, move, tag, level, 0, 0, w);
  return uniform53(w[0], w[1]);
}

// Polar Box-Muller pair (Common/Random/Random.h:156-180) on its own sub-stream.
__device__ inline void gaussian_pair(uint64_t seed, uint32_t stream, uint32_t move, int level, uint32_t pair,
                                     double& g0, double& g1) {
  for (uint32_t a = 0; a < 0x10000u; ++a) {
    uint32_t w[4];
    rng_block(seed, stream, move, kTagGauss, level, a, pair, w);
    double V1 = 2.0 * uniform53(w[0], w[1]) - 1.0;
    double V2 = 2.0 * uniform53(w[2], w[3]) - 1.0;
    double S = V1 * V1 + V2 * V2;
    if (S < 1.0 && S > 0.0) {
      double temp = sqrt((-2.0 * log(S)) / S);
      g0 = V1 * temp;
      g1 = V2 * temp;
      return;
    }
  }
  g0 = g1 = 0.0;
}

// --------------------------------------------------------------------------
// Periodic geometry (PIMC++/src/PathClass.h:696-746, 897-921; PathClass.cc:1389-1407)
// --------------------------------------------------------------------------

// PutInBox: n = -floor(v/L + 1/2); v += n L  (ties go up).
__device__ inline double put_in_box_1(double v, double box, double boxinv, double isper) {
#pragma clang fp contract(off)
  double n = -floor(v * boxinv + 0.5);
  return v + n * isper * box;
}

// PutInBoxFast.
__device__ inline double put_in_box_fast_1(double v, double box) {
#pragma clang fp contract(off)
  double half = box / 2;
  if (v > half) v -= box;
  if (v < -half) v += box;
  return v;
}

// DistDisp(sliceA, sliceB, p1, p2): dispA = r2(A) - r1(A) moved to its nearest
// image n = -nearbyint(dispA/L); dispB = r2(B) - r1(B) moved by
// m = nearbyint((dispA - dispB)/L), the image consistent with dispA.
__device__ inline void dist_disp_1(double a, double b, double box, double boxinv, double isper, double& da,
                                   double& db, double& n, double& m) {
#pragma clang fp contract(off)
  n = -rint(a * boxinv);
  da = a + n * isper * box;
  m = rint((da - b) * boxinv);
  db = b + m * isper * box;
}

// --------------------------------------------------------------------------
// Spline interval lookup (Common/Splines/Grid.h:81-84, 527-530; CubicSpline.h:490-501)
// --------------------------------------------------------------------------
__device__ inline int clamp_to_int(double t, int n) {
  t = fmax(t, -2.0);
  t = fmin(t, (double)(n + 1));
  return (int)t;
}

// Returns `low`, the interval index the reference's MultiCubicSpline uses.
// LINEAR: ReverseMap = (int)nearbyint((x-Start)*deltainv - 0.5).  LOG: the
// reference's floor(log(x/r0)/log(spacing)), corrected by one step against the
// stored knots so that a last-bit difference in log() cannot change it.
// LIN: the caller guarantees a LINEAR grid (no log-grid code is generated).
template <bool LIN = false, typename XPtr>
__device__ inline int spline_low(const FitDev& F, XPtr xs, double x) {
#pragma clang fp contract(off)
  int i;
  if (LIN || F.grid_type == 0) {
    i = clamp_to_int(rint((x - F.start) * F.deltainv - 0.5), F.ng);
  } else {
    i = clamp_to_int(floor(log(x / F.r0) / F.logspacing), F.ng);
    if (i >= 1 && i <= F.ng - 2 && x < xs[i])
      i -= 1;
    else if (i >= -1 && i <= F.ng - 3 && x >= xs[i + 1])
      i += 1;
  }
  int hi = i + 1;
  if (hi < 1) hi = 1;
  if (hi > F.ng - 1) hi = F.ng - 1;
  return hi - 1;
}

// The interval ix of x on G clamped to [0, n-2] (BicubicSpline's ix) and its
// record {X_i, h_i, 1/h_i, X_{i+1}}: an estimate from the grid's closed form (an
// FP32 log2 on a LOG grid), corrected by one step against the record's knots, so
// the interval is the one whose knots bracket x.  The estimate is within one
// interval of it for every grid these tables use (FP32 log2 errs by ~1e-7).
__device__ inline int grid_iv(const GridDev& G, double x, d4v& rec) {
  float t;
  if (G.type == 0)
    t = (float)((x - G.start) * G.deltainv);
  else
    t = __log2f((float)(x * G.r0inv)) * G.log2inv;
  t = fminf(fmaxf(t, 0.0f), (float)(G.n - 2));   // also maps log2(0) = -inf to 0
  int i = (int)t;
  const PIMC_GLOBAL d4v* iv = (const PIMC_GLOBAL d4v*)G.iv;   // global memory: plain vector loads
  rec = iv[i];
  if (x < rec.x && i > 0) {
    i -= 1;
    rec = iv[i];
  } else if (x >= rec.w && i < G.n - 2) {
    i += 1;
    rec = iv[i];
  }
  return i;
}

// --------------------------------------------------------------------------
// PAcoulombBCFitClass::U (PAcoulombBCFit.cc:159-176) with BicubicSpline
// evaluation (BicubicSpline.h:266-300, 356-413); the table is read from global
// memory.  The local coordinates are (x - X_i) times the stored 1/h_i.
// --------------------------------------------------------------------------
// The cell of (x, y): interval indices and records on the q and t grids.
struct BcCell {
  int ix, iy;
  d4v qr, tr;
};
__device__ inline BcCell bc_cell(const FitDev& F, const GridDev& Q, double x, double y) {
  BcCell c;
  c.ix = grid_iv(Q, x, c.qr);
  c.iy = grid_iv(F.tgrid, y, c.tr);
  return c;
}

__device__ inline double bicubic_at(const FitDev& F, const d4v* __restrict__ tab, const BcCell& c, double x,
                                    double y) {
  const int ix = c.ix, iy = c.iy;
  const double h = c.qr.y, k = c.tr.y;
  const double u = (x - c.qr.x) * c.qr.z;
  const double v = (y - c.tr.x) * c.tr.z;
  const double a0 = (u - 1.0) * (u - 1.0) * (1.0 + 2.0 * u), a1 = u * u * (3.0 - 2.0 * u);
  const double a2 = h * (u * (u - 1.0) * (u - 1.0)), a3 = h * (u * u * (u - 1.0));
  const double b0 = (v - 1.0) * (v - 1.0) * (1.0 + 2.0 * v), b1 = v * v * (3.0 - 2.0 * v);
  const double b2 = k * (v * (v - 1.0) * (v - 1.0)), b3 = k * (v * v * (v - 1.0));
  const PIMC_GLOBAL d4v* g = (const PIMC_GLOBAL d4v*)tab;
  const d4v A = g[(size_t)ix * F.nt + iy], B = g[(size_t)ix * F.nt + iy + 1];
  const d4v C = g[(size_t)(ix + 1) * F.nt + iy], D = g[(size_t)(ix + 1) * F.nt + iy + 1];
  // BCpoint = {z, dzdx, dzdy, d2zdxdy}
  return a0 * (A.x * b0 + B.x * b1 + A.z * b2 + B.z * b3) + a1 * (C.x * b0 + D.x * b1 + C.z * b2 + D.z * b3) +
         a2 * (A.y * b0 + B.y * b1 + A.w * b2 + B.w * b3) + a3 * (C.y * b0 + D.y * b1 + C.w * b2 + D.w * b3);
}

__device__ inline double bicubic(const FitDev& F, const GridDev& Q, const d4v* __restrict__ tab, double x,
                                 double y) {
  return bicubic_at(F, tab, bc_cell(F, Q, x, y), x, y);
}

// The same cell in the mixed path, anchored on its first corner: with A the node value
// z at (ix, iy), U = A + sum_ij a_i(u) b_j(v) (z_ij - A) + (the derivative terms), which is
// the same interpolant because a0 + a1 = b0 + b1 = 1.  The corner differences are formed
// in FP64 from the FP64 nodes; the basis and the sum run in FP32 on the differences and on
// the FP32 node derivatives tab32 = {dz/dx, dz/dy, d2z/dxdy, 0}, so the FP32 rounding is
// relative to the variation across the cell, not to U.
__device__ inline double bicubic32_at(const FitDev& F, const d4v* __restrict__ tab, const f4v* __restrict__ tab32,
                                      const BcCell& c, double x, double y) {
  const int ix = c.ix, iy = c.iy;
  const float h = (float)c.qr.y, k = (float)c.tr.y;
  const float u = (float)((x - c.qr.x) * c.qr.z);
  const float v = (float)((y - c.tr.x) * c.tr.z);
  const float a0 = (u - 1.0f) * (u - 1.0f) * (1.0f + 2.0f * u), a1 = u * u * (3.0f - 2.0f * u);
  const float a2 = h * (u * (u - 1.0f) * (u - 1.0f)), a3 = h * (u * u * (u - 1.0f));
  const float b0 = (v - 1.0f) * (v - 1.0f) * (1.0f + 2.0f * v), b1 = v * v * (3.0f - 2.0f * v);
  const float b2 = k * (v * (v - 1.0f) * (v - 1.0f)), b3 = k * (v * v * (v - 1.0f));
  const size_t iA = (size_t)ix * F.nt + iy, iC = (size_t)(ix + 1) * F.nt + iy;
  const PIMC_GLOBAL double* z = (const PIMC_GLOBAL double*)tab;   // z is the first field of a BCpoint
  const double zA = z[4 * iA], zB = z[4 * (iA + 1)], zC = z[4 * iC], zD = z[4 * (iC + 1)];
  const float dB = (float)(zB - zA), dC = (float)(zC - zA), dD = (float)(zD - zA);
  const PIMC_GLOBAL f4v* g = (const PIMC_GLOBAL f4v*)tab32;
  const f4v A = g[iA], B = g[iA + 1], C = g[iC], D = g[iC + 1];
  const float s = a0 * (dB * b1 + A.y * b2 + B.y * b3) + a1 * (dC * b0 + dD * b1 + C.y * b2 + D.y * b3) +
                  a2 * (A.x * b0 + B.x * b1 + A.z * b2 + B.z * b3) + a3 * (C.x * b0 + D.x * b1 + C.z * b2 + D.z * b3);
  return zA + (double)s;
}

__device__ inline double pair_u_bc(const FitDev& F, const GridDev& Q, int level, double q, double s2) {
  const d4v* tab = F.bc + (size_t)level * F.ng * F.nt;
  if (q <= (F.end * 1.0000001)) {
    if (q == 0) return bicubic(F, Q, tab, 0.0, 0.0);
    const double t = sqrt(s2) / (2.0 * q);
    return bicubic(F, Q, tab, q, t);
  }
  double beta = F.smallest_beta;
  for (int i = 0; i < level; i++) beta *= 2.0;
  return beta * (F.z1z2 / q);
}

// Mixed-precision PAcoulombBCFitClass::U: the cells of q intervals at and above i32 (the
// level's bound) in FP32; in FP64 the cells below it (near the core), q below the first
// knot (an extrapolation, where the Hermite basis grows large and cancels) and the
// Coulomb tail.
__device__ inline double pair_u_bc_mixed(const FitDev& F, const GridDev& Q, int level, int i32, double q,
                                         double s2) {
  if (q <= (F.end * 1.0000001)) {
    const double x = q, y = q == 0 ? 0.0 : sqrt(s2) / (2.0 * q);
    const BcCell c = bc_cell(F, Q, x, y);
    const size_t lev = (size_t)level * F.ng * F.nt;
    return (c.ix < i32 || x < c.qr.x) ? bicubic_at(F, F.bc + lev, c, x, y)
                                      : bicubic32_at(F, F.bc + lev, (const f4v*)F.c32 + lev, c, x, y);
  }
  double beta = F.smallest_beta;
  for (int i = 0; i < level; i++) beta *= 2.0;
  return beta * (F.z1z2 / q);
}

// QuinticSpline::operator() (QuinticSpline.h:172-218) of the replica's Ulong table.
// The knot I of x and P = x - X_I.
__device__ inline int ulong_knot(const RepCore& R, double x, double& P) {
  int J;
  {
#pragma clang fp contract(off)
    J = clamp_to_int(rint((x - R.ulong_start) * R.ulong_deltainv - 0.5), R.ulong_n) + 1;
  }
  int I = J - 1;
  if (I < 0) {
    I = 0;
    J = 1;
  }
  if (J > (R.ulong_n - 1)) {
    J = R.ulong_n - 1;
    I = J - 1;
  }
  double X;
  {
#pragma clang fp contract(off)
    X = R.ulong_start + (double)I * R.ulong_delta;
  }
  P = (x - X);
  return I;
}
__device__ inline double ulong_eval(const RepCore& R, const double* __restrict__ tab, double x) {
  double P;
  const int I = ulong_knot(R, x, P);
  const PIMC_GLOBAL double* k = (const PIMC_GLOBAL double*)tab + (size_t)I * 6;
  return ((((k[5] * P + k[4]) * P + k[3]) * P + k[2]) * P + k[1]) * P + k[0];
}
// Mixed path: the knot value Y in FP64 plus the rest of the same Horner form,
// (((F P + E) P + D) P + C) P + B) P, in FP32 (P is at most one grid spacing, so this part
// is small next to Y); one 32-byte record per knot.
__device__ inline double ulong_eval32(const RepCore& R, const UlongKnot32* __restrict__ tab32, double x) {
  double P;
  const int I = ulong_knot(R, x, P);
  const PIMC_GLOBAL UlongKnot32* k = (const PIMC_GLOBAL UlongKnot32*)tab32 + I;
  const float p = (float)P;
  return k->y + (double)(((((k->c[4] * p + k->c[3]) * p + k->c[2]) * p + k->c[1]) * p + k->c[0]) * p);
}

// --------------------------------------------------------------------------
// Pair action U(q, z, s^2) of a DavidPAClass fit at one tau level
// (Common/PairAction/DavidPAClass.cc:92-102, 468-511), tables in LDS.
// --------------------------------------------------------------------------
template <bool InLds>
struct LevelTable64 {
  typename Ptr<InLds>::template c<d2v> geo;   // [nint]
  typename Ptr<InLds>::template c<d4v> c;     // [nint][ncol]
  typename Ptr<InLds>::template c<double> x;  // [ng] (log grid correction)
};
template <bool InLds>
struct LevelTable32 {
  typename Ptr<InLds>::template c<d2v> geo;
  typename Ptr<InLds>::template c<f4v> c;            // FP32 coefficients
  typename Ptr<InLds>::template c<double> x;
  const d4v* c64;                                     // FP64 coefficients of this level (global memory)
  // staged tables: the FP64 rows of the first n64s intervals copied into LDS (n64s is
  // FitDev::i32_lds: every interval the mixed lookups read in FP64, or 0)
  const PIMC_LDS d4v* c64s;
  int n64s;
  int i32;   // FitDev::i32_lev of this level: intervals below it are evaluated from the FP64 rows
};

__device__ inline double horner64(const d4v c, double b) { return c.x + b * (c.y + b * (c.z + b * c.w)); }
__device__ inline float horner32(const f4v c, float b) { return c.x + b * (c.y + b * (c.z + b * c.w)); }

template <int NT, bool InLds>
__device__ inline double pair_u64(const FitDev& F, const LevelTable64<InLds>& T, double q, double z, double s2) {
  const int ncol = F.ncol;
  double rmin = F.start;
  double r = q + 0.5 * z;
  double rp = q - 0.5 * z;
  if (r > F.end) r = F.end;
  if (rp > F.end) rp = F.end;
  if ((rp < rmin) || (r < rmin)) return T.c[0].x;  // u00(rmin): interval 0 at b = 0
  int lr = spline_low(F, T.x, r);
  int lp = spline_low(F, T.x, rp);
  d2v gr = T.geo[lr], gp = T.geo[lp];
  double br = (r - gr.x) * gr.y;
  double bp = (rp - gp.x) * gp.y;
  double U = 0.5 * (horner64(T.c[lr * ncol], br) + horner64(T.c[lp * ncol], bp));
  if constexpr (NT > 0) {
    if (!(s2 > 0.0 && q < F.end)) return U;
    int lq = spline_low(F, T.x, q);
    d2v gq = T.geo[lq];
    double bq = (q - gq.x) * gq.y;
    const int cq = lq * ncol + 1;
    // sum_{k=1..NT} sum_{j=0..k} u_kj(q) z^{2j} s^{2(k-j)}, table column k(k+1)/2 + j + 1
    double zp[NT + 1], sp[NT + 1];
    zp[0] = 1.0;
    sp[0] = 1.0;
#pragma unroll
    for (int t = 1; t <= NT; t++) {
      zp[t] = zp[t - 1] * (z * z);
      sp[t] = sp[t - 1] * s2;
    }
    double off = 0.0;
    int col = 0;
#pragma unroll
    for (int k = 1; k <= NT; k++)
#pragma unroll
      for (int j = 0; j <= k; j++, col++) off += horner64(T.c[cq + col], bq) * (zp[j] * sp[k - j]);
    U += off;
  }
  return U;
}

// sqrt in FP64 from an FP32 reciprocal-square-root seed and one Goldschmidt
// step (relative error ~1e-15); used by the mixed path for pair distances.
// Valid for 0 <= a < 1e30 (squared pair distances): a = 0 gives 0, and a below
// 1e-30 (distances below 1e-15, inside any pair table's core) is not refined.
__device__ inline double sqrt_seeded(double a) {
  double y = (double)__builtin_amdgcn_rsqf(fmaxf((float)a, 1e-30f));
  double g = a * y;
  double h = 0.5 * y;
  double d = fma(-g, h, 0.5);
  return fma(g, d, g);
}

// Columns [c0col, c0col + NC) of interval i at local coordinate b: FP32 Horner
// from the FP32 table, or FP64 from the global FP64 table when use64 (the caller
// decides from the lookup coordinate, before any table load).
template <int NC, bool InLds>
__device__ inline void eval_cols_split(const LevelTable32<InLds>& T, int ncol, int i, int c0col, double b, bool use64,
                                       double out[NC]) {
  const int at = i * ncol + c0col;
  if (use64) {
    if (InLds && T.n64s > 0) {
      const PIMC_LDS d4v* c64 = T.c64s + (size_t)at;
#pragma unroll
      for (int k = 0; k < NC; k++) out[k] = horner64(c64[k], b);
    } else {
      const d4v* c64 = T.c64 + (size_t)at;
#pragma unroll
      for (int k = 0; k < NC; k++) out[k] = horner64(c64[k], b);
    }
  } else {
    const float bf = (float)b;
#pragma unroll
    for (int k = 0; k < NC; k++) out[k] = (double)horner32(T.c[at + k], bf);
  }
}

// Columns [c0col, c0col + NC) of interval i at local coordinate b: FP32 Horner
// from the FP32 table, or FP64 from the global FP64 table where the interval is
// flagged (its FP32 row holds NaN).
template <int NC, bool InLds>
__device__ inline void eval_cols_mixed(const LevelTable32<InLds>& T, int ncol, int i, int c0col, double b,
                                       double out[NC]) {
  const int at = i * ncol + c0col;
  f4v c[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) c[k] = T.c[at + k];
  if (__builtin_isnan(c[0].x)) {
    const d4v* c64 = T.c64 + (size_t)at;
#pragma unroll
    for (int k = 0; k < NC; k++) out[k] = horner64(c64[k], b);
  } else {
    const float bf = (float)b;
#pragma unroll
    for (int k = 0; k < NC; k++) out[k] = (double)horner32(c[k], bf);
  }
}

// Mixed-precision U: FP64 geometry, interval lookup and powers; the spline
// polynomials in FP32 except on flagged intervals; the sum in FP64.
template <int NT, bool InLds>
__device__ inline double pair_u_mixed(const FitDev& F, const LevelTable32<InLds>& T, double q, double z, double s2) {
  const int ncol = F.ncol;
  double rmin = F.start;
  double r = q + 0.5 * z;
  double rp = q - 0.5 * z;
  if (r > F.end) r = F.end;
  if (rp > F.end) rp = F.end;
  if ((rp < rmin) || (r < rmin)) return T.c64[0].x;  // u00(rmin)
  int lr = spline_low(F, T.x, r);
  int lp = spline_low(F, T.x, rp);
  d2v gr = T.geo[lr], gp = T.geo[lp];
  double ur[1], up[1];
  eval_cols_mixed<1>(T, ncol, lr, 0, (r - gr.x) * gr.y, ur);
  eval_cols_mixed<1>(T, ncol, lp, 0, (rp - gp.x) * gp.y, up);
  double U = 0.5 * (ur[0] + up[0]);
  if constexpr (NT > 0) {
    if (!(s2 > 0.0 && q < F.end)) return U;
    constexpr int NOFF = NT * (NT + 3) / 2;
    int lq = spline_low(F, T.x, q);
    d2v gq = T.geo[lq];
    double cv[NOFF];
    eval_cols_mixed<NOFF>(T, ncol, lq, 1, (q - gq.x) * gq.y, cv);
    double zp[NT + 1], sp[NT + 1];
    zp[0] = 1.0;
    sp[0] = 1.0;
#pragma unroll
    for (int t = 1; t <= NT; t++) {
      zp[t] = zp[t - 1] * (z * z);
      sp[t] = sp[t - 1] * s2;
    }
    double off = 0.0;
    int col = 0;
#pragma unroll
    for (int k = 1; k <= NT; k++)
#pragma unroll
      for (int j = 0; j <= k; j++, col++) off += cv[col] * (zp[j] * sp[k - j]);
    U += off;
  }
  return U;
}

}  // namespace pimc
