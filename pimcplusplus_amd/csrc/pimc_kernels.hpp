// pimc_kernels.hpp -- kernel building blocks of the action library (gfx950, wave64).
#pragma once

#include "pimc_device.hpp"
#include "../../include/pimc_gpu.h"

namespace pimc {

constexpr int kBlock = 256;          // threads per block: 4 waves
constexpr int kWaves = kBlock / 64;
constexpr int kMaxWaves = 8;         // largest block of the bisection kernel: 512 threads

// Request of the SingleAction entry points, as uploaded to the device.
struct ReqDev {
  int replica, mode, slice1, slice2, level, nptcls;
  int ptcls[kMaxActive];
};

// A block-uniform copy of *p: every 32-bit word goes through readfirstlane, so
// the copy lives in scalar registers instead of one copy per lane.
template <typename T>
__device__ inline T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized struct");
  T out;
  const int* src = (const int*)p;
  int* dst = (int*)&out;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); i++) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

// --------------------------------------------------------------------------
// Reductions
// --------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of v over the block, in a fixed order; the result is valid in thread 0.
__device__ inline double block_sum(double v, PIMC_LDS double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += scratch[w];
  __syncthreads();
  return t;
}

// Block sums of two values in one pass (valid in thread 0), in a fixed order.  Without the
// trailing barrier the caller must order the next write of scratch after thread 0's reads.
__device__ inline void block_sum2(double a, double b, PIMC_LDS double* scratch, double& ta, double& tb,
                                  bool trailing_barrier = true) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  if (lane == 0) {
    scratch[wave] = a;
    scratch[nw + wave] = b;
  }
  __syncthreads();
  ta = 0.0;
  tb = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; w++) {
      ta += scratch[w];
      tb += scratch[nw + w];
    }
  if (trailing_barrier) __syncthreads();
}

// Kahan-compensated FP64 accumulation of the mixed path's terms.
struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double x) {
    double y = x - c;
    double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline double value() const { return s; }
};

// One thread's short-range sums of a stage.  The FP64 path sums the OLD and NEW link actions
// apart, as the reference does (a = OLD, b = NEW).  The mixed path sums OLD plainly and the
// change NEW - OLD of every link pair with Kahan compensation (a = OLD, b = the change): the
// change is what the Metropolis test uses, and one compensated sum per pair costs half of two.
template <bool MIXED>
struct StageSums {
  double a = 0.0, b = 0.0, c = 0.0;
  __device__ inline void add(double uo, double un) {
    a += uo;
    if constexpr (MIXED) {
      const double y = (un - uo) - c;
      const double t = b + y;
      c = (t - b) - y;
      b = t;
    } else {
      b += un;
    }
  }
  // NEW total from the block totals of a and b
  __device__ static inline double new_total(double ta, double tb) { return MIXED ? ta + tb : tb; }
  // b's contribution of a pair of terms added outside add()
  __device__ static inline double b_of(double uo, double un) { return MIXED ? un - uo : un; }
};

// --------------------------------------------------------------------------
// Table views: one tau level of a fit, in global memory or staged in LDS
// --------------------------------------------------------------------------
__device__ inline LevelTable64<false> level_table64(const FitDev& F, int lev) {
  LevelTable64<false> t;
  t.geo = F.geo;
  t.c = F.c64 + (size_t)lev * F.nint * F.ncol;
  t.x = F.x;
  return t;
}
__device__ inline LevelTable32<false> level_table32(const FitDev& F, int lev) {
  LevelTable32<false> t;
  t.geo = F.geo;
  t.c = F.c32 + (size_t)lev * F.nint * F.ncol;
  t.x = F.x;
  t.c64 = F.c64 + (size_t)lev * F.nint * F.ncol;
  t.c64s = nullptr;
  t.n64s = 0;
  t.i32 = F.i32_lev[lev];
  return t;
}

template <int NT, bool InLds>
__device__ inline double pair_u(const FitDev& F, const LevelTable64<InLds>& T, double q, double z, double s2) {
  return pair_u64<NT>(F, T, q, z, s2);
}
template <int NT, bool InLds>
__device__ inline double pair_u(const FitDev& F, const LevelTable32<InLds>& T, double q, double z, double s2) {
  return pair_u_mixed<NT>(F, T, q, z, s2);
}

// A level of a PAcoulombBCFitClass fit (tables in global memory); i32 is the level's bound
// below which the mixed path evaluates cells from the FP64 nodes.
struct BcView {
  GridDev q;
  int level;
  int i32;
};
// The same level evaluated by the mixed path.
struct BcView32 : BcView {};
__device__ inline GridDev qgrid_of(const FitDev& F) { return F.qgrid; }
template <int NT>
__device__ inline double pair_u(const FitDev& F, const BcView& T, double q, double z, double s2) {
  return pair_u_bc(F, T.q, T.level, q, s2);
}
template <int NT>
__device__ inline double pair_u(const FitDev& F, const BcView32& T, double q, double z, double s2) {
  return pair_u_bc_mixed(F, T.q, T.level, T.i32, q, s2);
}

template <typename Table>
struct IsMixed {
  static constexpr bool value = false;
};
template <bool InLds>
struct IsMixed<LevelTable32<InLds>> {
  static constexpr bool value = true;
};
template <>
struct IsMixed<BcView32> {
  static constexpr bool value = true;
};

// The Ulong tables of one species pair and level: the FP64 knots and their B..F in FP32
// (mixed path); t64 is nullptr where ShortRange subtracts no Ulong.
struct UlongView {
  const double* t64;
  const UlongKnot32* t32;
};
template <bool MIXED>
__device__ inline double ulong_at(const RepCore& R, const UlongView& u, double x) {
  if constexpr (MIXED)
    return ulong_eval32(R, u.t32, x);
  else
    return ulong_eval(R, u.t64, x);
}

// Bytes of LDS one staged level of fit F takes: geo + coefficients + grid, or the
// coefficients alone when the lookups use lin_low (lin); a mixed stage adds the FP64
// rows of its first n64s intervals, 32-byte aligned.
__host__ __device__ inline size_t staged32_fp64_offset(int nint, int ncol, int ng, bool lin) {
  const size_t b = (size_t)nint * ncol * 16 + (lin ? 0 : (size_t)nint * 16 + (size_t)ng * 8);
  return (b + 31) & ~(size_t)31;
}
__host__ __device__ inline size_t staged_bytes(int nint, int ncol, int ng, bool mixed, bool lin = false,
                                               int n64s = 0) {
  if (mixed) return staged32_fp64_offset(nint, ncol, ng, lin) + (size_t)n64s * ncol * 32;
  const size_t coef = (size_t)nint * ncol * 32;
  const size_t b = lin ? coef : (size_t)nint * 16 + coef + (size_t)ng * 8;
  return (b + 15) & ~(size_t)15;
}

// Block-wide copy of n elements from global memory to LDS, four loads in flight per thread.
template <typename T>
__device__ inline void block_copy(PIMC_LDS T* dst, const T* __restrict__ src, int n) {
  const int bs = blockDim.x;
  for (int base = threadIdx.x; base < n; base += 4 * bs) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (base + u * bs < n) v[u] = src[base + u * bs];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (base + u * bs < n) dst[base + u * bs] = v[u];
  }
}

// Cooperative copy of level `lev` of F into LDS at `base` (FP64 view).  LIN:
// the lookups use lin_low, so the interval geometry and knots are not copied.
template <bool LIN = false>
__device__ inline LevelTable64<true> stage_level64(const FitDev& F, int lev, PIMC_LDS unsigned char* base) {
  PIMC_LDS d2v* geo = (PIMC_LDS d2v*)base;
  PIMC_LDS d4v* c = (PIMC_LDS d4v*)(LIN ? base : base + (size_t)F.nint * 16);
  const int nc = F.nint * F.ncol;
  PIMC_LDS double* xs = (PIMC_LDS double*)((PIMC_LDS unsigned char*)c + (size_t)nc * 32);
  if (!LIN) block_copy(geo, F.geo, F.nint);
  block_copy(c, F.c64 + (size_t)lev * nc, nc);
  if (!LIN) block_copy(xs, F.x, F.ng);
  LevelTable64<true> t;
  t.geo = geo;
  t.c = c;
  t.x = xs;
  return t;
}

// Cooperative copy of level `lev` of F into LDS at `base` (mixed view).
template <bool LIN = false>
__device__ inline LevelTable32<true> stage_level32(const FitDev& F, int lev, PIMC_LDS unsigned char* base) {
  PIMC_LDS d2v* geo = (PIMC_LDS d2v*)base;
  PIMC_LDS f4v* c = (PIMC_LDS f4v*)(LIN ? base : base + (size_t)F.nint * 16);
  const int nc = F.nint * F.ncol;
  PIMC_LDS double* xs = (PIMC_LDS double*)((PIMC_LDS unsigned char*)c + (size_t)nc * 16);
  PIMC_LDS d4v* c64s = (PIMC_LDS d4v*)(base + staged32_fp64_offset(F.nint, F.ncol, F.ng, LIN));
  if (!LIN) block_copy(geo, F.geo, F.nint);
  block_copy(c, F.c32 + (size_t)lev * nc, nc);
  if (!LIN) block_copy(xs, F.x, F.ng);
  block_copy(c64s, F.c64 + (size_t)lev * nc, F.i32_lds * F.ncol);
  LevelTable32<true> t;
  t.geo = geo;
  t.c = c;
  t.x = xs;
  t.c64 = F.c64 + (size_t)lev * nc;
  t.c64s = c64s;
  t.n64s = F.i32_lds;
  t.i32 = F.i32_lev[lev];
  return t;
}

// --------------------------------------------------------------------------
// One pair-link term: DistDisp + U for moved bead positions (m_a, m_b) and
// partner positions (p_a, p_b), returning U (ShortRangeClass.cc:153-167).
// --------------------------------------------------------------------------
template <int NT, typename Table>
__device__ inline double pair_link_u(const RepCore& R, const FitDev& F, const Table& T, const double ma[3],
                                     const double mb[3], const double pa[3], const double pb[3],
                                     const UlongView ul = UlongView{nullptr, nullptr}) {
  constexpr bool kMixed = IsMixed<Table>::value;
  double da[3], db[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double n, m;
    dist_disp_1(pa[d] - ma[d], pb[d] - mb[d], R.box[d], R.boxinv[d], R.isper[d], da[d], db[d], n, m);
  }
  double a2 = da[0] * da[0] + da[1] * da[1] + da[2] * da[2];
  double b2 = db[0] * db[0] + db[1] * db[1] + db[2] * db[2];
  double ra = kMixed ? sqrt_seeded(a2) : sqrt(a2);
  double rb = kMixed ? sqrt_seeded(b2) : sqrt(b2);
  double e0 = da[0] - db[0], e1 = da[1] - db[1], e2 = da[2] - db[2];
  double s2 = e0 * e0 + e1 * e1 + e2 * e2;
  double q = 0.5 * (ra + rb);
  double z = ra - rb;
  double U = pair_u<NT>(F, T, q, z, s2);
  // long-ranged fits: subtract the part the k-space action carries (ShortRangeClass.cc:165-166)
  if (ul.t64) U -= 0.5 * (ulong_at<kMixed>(R, ul, ra) + ulong_at<kMixed>(R, ul, rb));
  return U;
}

// --------------------------------------------------------------------------
// calcUsqzFast split by what it depends on (DavidPAClass.cc:468-511), for loops
// that walk a partner's links in order: the diagonal column u00 depends on one
// slice only and is evaluated once per slice, at |r| clamped to End (the
// reference evaluates it at q +- z/2, the same distance up to rounding); the
// off-diagonal sum depends on the link.
// --------------------------------------------------------------------------
// Interval and local coordinate of x on a LINEAR grid: t = (x - Start)/delta,
// i = clamp(floor t) and b = t - i, for 0 <= x <= End.  The reference's
// ReverseMap (nearbyint(t - 1/2)) differs only where t is an integer, where both
// neighbouring cubics take the same value; b differs from (x - X_i)/h by
// rounding of t.
// x < Start (a distance inside the core, whose value the caller replaces by
// u00(rmin)) maps to an in-range interval through the unsigned clamp.
__device__ inline int lin_low(const FitDev& F, double x, double& b) {
  const double t = (x - F.start) * F.deltainv;
  const int i = (int)min((unsigned)(int)t, (unsigned)(F.nint - 1));
  b = t - (double)i;
  return i;
}

// Interval and local coordinate of x by the grid's own lookup.
template <bool LIN, typename Table>
__device__ inline int table_low(const FitDev& F, const Table& T, double x, double& b) {
  if constexpr (LIN) {
    return lin_low(F, x, b);
  } else {
    const int l = spline_low(F, T.x, x);
    const d2v g = T.geo[l];
    b = (x - g.x) * g.y;
    return l;
  }
}

template <bool LIN, bool InLds>
__device__ inline double slice_u00(const FitDev& F, const LevelTable64<InLds>& T, double r) {
  r = fmin(r, F.end);
  double b;
  const int l = table_low<LIN>(F, T, r, b);
  return horner64(T.c[l * F.ncol], b);
}
template <bool LIN, bool InLds>
__device__ inline double slice_u00(const FitDev& F, const LevelTable32<InLds>& T, double r) {
  r = fmin(r, F.end);
  double b;
  const int l = table_low<LIN>(F, T, r, b);
  double u[1];
  eval_cols_split<1>(T, F.ncol, l, 0, b, l < T.i32, u);
  return u[0];
}

// sqrt for the short-range distances (0 <= a < 1e300): the hardware reciprocal
// square root refined by one Goldschmidt and two Newton steps, without the
// denormal scaling of the library sqrt; a = 0 gives ~1e-150.
__device__ inline double sqrt_dist(double a) {
  a = fmax(a, 1e-300);
  const double y = __builtin_amdgcn_rsq(a);
  double g = a * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, a);
  g = fma(d, h, g);
  d = fma(-g, g, a);
  return fma(d, h, g);
}

// sum_{k=1..NT} sum_{j=0..k} u_kj(q) z^{2j} s^{2(k-j)} where s > 0 and q < End, else 0.
template <int NT, bool LIN, typename Table>
__device__ inline double link_offdiag(const FitDev& F, const Table& T, double q, double z, double s2) {
  if constexpr (NT == 0) {
    return 0.0;
  } else {
    if (!(s2 > 0.0 && q < F.end)) return 0.0;
    constexpr int NOFF = NT * (NT + 3) / 2;
    double bq;
    const int lq = table_low<LIN>(F, T, q, bq);
    double cv[NOFF];
    if constexpr (IsMixed<Table>::value) {
      eval_cols_split<NOFF>(T, F.ncol, lq, 1, bq, lq < T.i32, cv);
    } else {
#pragma unroll
      for (int c = 0; c < NOFF; c++) cv[c] = horner64(T.c[lq * F.ncol + 1 + c], bq);
    }
    const double z2 = z * z;
    if constexpr (NT == 1) {
      return fma(cv[1], z2, cv[0] * s2);   // u10 s^2 + u11 z^2
    } else if constexpr (NT == 2) {
      // u10 s^2 + u11 z^2 + u20 s^4 + u21 z^2 s^2 + u22 z^4, grouped in s^2 and z^2
      const double a = fma(cv[4], z2, cv[1]);
      const double c = fma(cv[3], z2, fma(cv[2], s2, cv[0]));
      return fma(a, z2, c * s2);
    }
    double zp[NT + 1], sp[NT + 1];
    zp[0] = 1.0;
    sp[0] = 1.0;
#pragma unroll
    for (int t = 1; t <= NT; t++) {
      zp[t] = zp[t - 1] * z2;
      sp[t] = sp[t - 1] * s2;
    }
    double off = 0.0;
    int col = 0;
#pragma unroll
    for (int k = 1; k <= NT; k++)
#pragma unroll
      for (int j = 0; j <= k; j++, col++) off += cv[col] * (zp[j] * sp[k - j]);
    return off;
  }
}

// u00(rmin): the value at the first knot (interval 0 at b = 0).
template <bool InLds>
__device__ inline double u00_rmin(const LevelTable64<InLds>& T) { return T.c[0].x; }
template <bool InLds>
__device__ inline double u00_rmin(const LevelTable32<InLds>& T) { return T.c64[0].x; }

// Nearest-image displacement of one dimension (DistDisp's n = -nearbyint(a/L)).
__device__ inline double near_image_1(double a, double box, double boxinv, double isper, double& n) {
#pragma clang fp contract(off)
  n = -rint(a * boxinv);
  return a + n * isper * box;
}
// DistDisp's image of the second slice, consistent with the imaged first-slice displacement da.
__device__ inline double cons_image_1(double da, double b, double box, double boxinv, double isper, double& m) {
#pragma clang fp contract(off)
  m = rint((da - b) * boxinv);
  return b + m * isper * box;
}

// Ulong tables of pair `pair` at `level` when ShortRange must subtract them, else {nullptr, nullptr}.
__device__ inline UlongView ulong_for(const RepCore& R, const FitDev& F, int pair, int level) {
  if (F.kind != 1 || !R.use_long_range) return UlongView{nullptr, nullptr};
  const size_t knot = ((size_t)pair * R.ulong_levels + level) * R.ulong_n;
  return UlongView{R.ulong + knot * 6, R.ulong32 + knot};
}

__device__ inline void load_bead(const RepCore& R, int mode, int slice, int p, double out[3]) {
  const double* base = R.path(mode) + (size_t)slice * 3 * R.N;
  out[0] = base[p];
  out[1] = base[R.N + p];
  out[2] = base[2 * R.N + p];
}
__device__ inline void store_bead(const RepCore& R, int mode, int slice, int p, const double v[3]) {
  double* base = R.path(mode) + (size_t)slice * 3 * R.N;
  base[p] = v[0];
  base[R.N + p] = v[1];
  base[2 * R.N + p] = v[2];
}

// --------------------------------------------------------------------------
// Kinetic link action (KineticClass.cc:33-90): -log prod_d sum_images exp(-(v_d + iL)^2/(4 lambda tau))
// --------------------------------------------------------------------------
template <typename P>
__device__ inline double kinetic_link(const RepCore& R, P a, P b, double FourLambdaTauInv) {
  double GaussProd = 1.0;
#pragma unroll
  for (int dim = 0; dim < 3; dim++) {
    double vel = put_in_box_1(b[dim] - a[dim], R.box[dim], R.boxinv[dim], R.isper[dim]);
    double GaussSum = 0.0;
    for (int image = -R.num_images; image <= R.num_images; image++) {
      double dist = vel + (double)image * R.box[dim];
      GaussSum += exp(-dist * dist * FourLambdaTauInv);
    }
    GaussProd *= GaussSum;
  }
  return -log(GaussProd);
}

// --------------------------------------------------------------------------
// Structure factors (PathClass.cc:616-707): C_d(n) = e^{i n phi_d} for
// n in [-maxk_d, maxk_d] by the reference's recursion, and the k term
// C_0(kx) C_1(ky) C_2(kz).  T = double for FP64, float for the mixed phase sums.
// --------------------------------------------------------------------------
__device__ inline int kstride(const RepCore& R) { return 2 * (R.maxk0 + R.maxk1 + R.maxk2) + 3; }
// The fields of dimension d (0, 1, 2) are combined arithmetically rather than selected:
// the compiler turns a select between fields of the register-resident R into a load at a
// selected offset, which moves the whole of R into scratch memory when d varies by lane.
__device__ inline int kmax_of(const RepCore& R, int d) {
  return R.maxk0 * (d == 0) + R.maxk1 * (d == 1) + R.maxk2 * (d == 2);
}
__device__ inline int koff_of(const RepCore& R, int d) {
  return (2 * R.maxk0 + 1) * (d >= 1) + (2 * R.maxk1 + 1) * (d == 2);
}

// A phase-table entry (re, im) in the arithmetic of the phase sums: FP64 (d2v) or FP32 (f2v).
template <typename T>
using PhaseRow = typename std::conditional<sizeof(T) == 8, d2v, f2v>::type;

// cos and sin of phi = r k_box are taken as cospi/sinpi of phi/pi: |phi| stays
// within a few pi here, and the pi-scaled forms need no long argument reduction
// (whose tables and registers would otherwise be live inside the chain kernel).
template <typename T>
__device__ inline void phase_row(double phi, int maxk, PIMC_LDS PhaseRow<T>* C) {
  T c, s;
  constexpr double kInvPi = 0.31830988618379067154;
  if constexpr (sizeof(T) == 8) {
    sincospi(phi * kInvPi, &s, &c);
  } else {
    sincospif((float)(phi * kInvPi), &s, &c);
  }
  T re = 1, im = 0;
  C[maxk] = PhaseRow<T>{1, 0};
  for (int n = 1; n <= maxk; n++) {
    T nre = c * re - s * im;   // tempC * C(n-1)
    T nim = c * im + s * re;
    re = nre;
    im = nim;
    C[maxk + n] = PhaseRow<T>{re, im};
    C[maxk - n] = PhaseRow<T>{re, -im};
  }
}

// The three rows of one position r (dimension d at koff_of(d)); lanes 0..2 of the caller.
// d is a per-lane value: the dimension's fields are combined arithmetically, not indexed
// or selected (either would move the register-resident R to scratch).
template <typename T>
__device__ inline void phase_rows_lane(const RepCore& R, const double r[3], int d, PIMC_LDS PhaseRow<T>* C) {
  const double x = d == 0 ? r[0] : (d == 1 ? r[1] : r[2]);
  const double kb = R.kbox[0] * (double)(d == 0) + R.kbox[1] * (double)(d == 1) + R.kbox[2] * (double)(d == 2);
  phase_row<T>(x * kb, kmax_of(R, d), C + koff_of(R, d));
}

template <typename T>
__device__ inline void kterm(const PIMC_LDS PhaseRow<T>* C, int o1, int o2, const int* ki, T& re, T& im) {
  const PhaseRow<T> a = C[ki[0]], b = C[o1 + ki[1]], c = C[o2 + ki[2]];
  const T abr = a.x * b.x - a.y * b.y, abi = a.x * b.y + a.y * b.x;   // C0 * C1
  re = abr * c.x - abi * c.y;                                         // (C0 * C1) * C2
  im = abr * c.y + abi * c.x;
}

// The k-independent terms (Madelung, background, k=0) of one slice, added to
// homo / background / k0 as the reference orders them.
__device__ inline void lr_const_terms(const RepCore& R, const RepSpecies& S, int level, double factor, double& homo,
                                      double& background, double& k0) {
  const int L = R.lr_levels;
  for (int s = 0; s < R.ns; s++) {
    const int pr = s * R.ns + s;
    if (!((R.lr_mask >> pr) & 1u)) continue;  // the reference adds these for every species; tables are zero
    const int N = S.count[s];
    homo -= factor * 0.5 * N * R.lr_const[(size_t)(pr * 3 + 0) * L + level];
    background -= factor * 0.5 * N * N * R.lr_const[(size_t)(pr * 3 + 1) * L + level];
    k0 += factor * 0.5 * N * N * R.lr_const[(size_t)(pr * 3 + 2) * L + level];
  }
  for (int s1 = 0; s1 < R.ns; s1++)
    for (int s2 = s1 + 1; s2 < R.ns; s2++) {
      const int pr = s1 * R.ns + s2;
      if (!((R.lr_mask >> pr) & 1u)) continue;
      const int N1 = S.count[s1], N2 = S.count[s2];
      background -= factor * N1 * N2 * R.lr_const[(size_t)(pr * 3 + 1) * L + level];
      k0 += factor * N1 * N2 * R.lr_const[(size_t)(pr * 3 + 2) * L + level];
    }
}

}  // namespace pimc
