// pimc_oracle.cc -- TEST INFRASTRUCTURE ONLY (the parity oracle).
//
// A plain C++17 CPU restatement of PIMC++'s action hot path, in FP64, written
// to follow the reference's arithmetic statement by statement.  It implements
// every entry point of include/pimc_gpu.h so the tests can drive it and the
// GPU library through one interface, plus the hooks of oracle/pimc_oracle.h.
// It is compiled with -ffp-contract=off so no multiply-add is fused that the
// reference source does not fuse.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it.
//
// Reference files followed (paths relative to the PIMC++ source tree):
//   Common/Splines/Grid.h:81-84,527-547        grids and ReverseMap
//   Common/Splines/CubicSpline.h:470-561       MultiCubicSpline evaluation
//   Common/Splines/CubicSpline.cc:77-128       MultiCubicSpline::Update
//   Common/PairAction/DavidPAClass.cc:92-102,468-511,755-773,859-868
//   PIMC++/src/PathClass.h:696-746,831-865,897-921   DistDisp, Velocity, PutInBox
//   PIMC++/src/PathClass.cc:281-292,534-586,616-723,727-899, 1389-1407
//   PIMC++/src/Actions/ShortRangeClass.cc:131-176
//   PIMC++/src/Actions/KineticClass.cc:33-90
//   PIMC++/src/Actions/LongRangeClass.cc:273-371
//   PIMC++/src/Moves/StageClass.cc:49-95,124-136, MultiStage.cc:42-110,
//   PIMC++/src/Moves/BisectionStage.cc:128-289, BisectionBlock.cc:66-104,310-451,
//   PIMC++/src/Moves/NoPermuteClass.cc:26-68, Common/Random/Random.h:45-46,156-180
//   PIMC++/src/Moves/PermuteTableClass.cc:122-150,172-187   ConstructHTable, CycleClass::Apply
//   Common/Splines/BicubicSpline.h:24-30,266-300,356-413, BicubicSpline.cc:20-154
//   Common/PairAction/PAcoulombBCFit.cc:159-192     Coulomb U and dU
//   Common/Splines/QuinticSpline.h:172-218          Ulong / dUlong / Vlong evaluation
//   PIMC++/src/Actions/KineticClass.cc:187-256, ShortRangeClass.cc:256-295,
//   PIMC++/src/Actions/LongRangeClass.cc:374-442, ActionsClass.cc:813-871   energy estimator
//   PIMC++/src/Actions/ShortRangePotClass.cc:27-47, LongRangePotClass.cc:32-85   potential
//   PIMC++/src/Observables/ObservableCorrelation.cc:156-221                 pair correlation
//
// How each part is pinned (tests/): Philox by the Random123 known-answer vectors; grids,
// cubic, bicubic and quintic splines by scipy interpolants and brute-force interval
// searches; DavidPA U and dU by a numpy restatement over scipy splines on the shipped
// testcoul.dm table; DistDisp/PutInBox by brute-force image search with exact ties;
// the short-range, kinetic and long-range actions by brute-force sums, closed forms and
// the rock-salt Madelung constant; the estimators by tau-derivative identities, the free-
// particle 3N/(2 beta) limit and an independent Ewald sum of the potential; the
// bookkeeping by byte equality with numpy index operations.  Parity unpinned: the
// reference's own SPRNG2 stream (not available), so the reference's own accept sequence.
//
// Deliberate, documented departures:
//   * LogGrid::ReverseMap: the reference's floor(log(r/r0)/log(spacing)) is
//     corrected by one step against the stored knots so that host and device
//     logarithms that differ in the last bit cannot pick different intervals.
//     It differs from the literal formula only within rounding distance of a
//     knot, where both neighbouring cubic pieces give the same value.
//   * The random stream: SPRNG2 is not available, so the clone's Local stream
//     is Philox4x32-10 with the counter layout documented at rng_ctr() below.
//     The polar Box-Muller method is kept, with its second variate used next,
//     but pairs are indexed per (move, level) instead of a function-static cache.

#include "../include/pimc_gpu.h"
#include "pimc_oracle.h"

#include <cmath>
#include <complex>
#include <cstdint>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <limits>
#include <vector>
#include <algorithm>

namespace {

typedef std::complex<double> cplx;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ------------------------------------------------------------------------
void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
  uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
  uint32_t k0 = k_in[0], k1 = k_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double uniform53(uint32_t w0, uint32_t w1) {
  uint64_t v = ((uint64_t)w1 << 32) | (uint64_t)w0;
  return (double)(v >> 11) * 0x1.0p-53;
}

enum { TAG_SLICE = 1, TAG_PTCL = 2, TAG_GAUSS = 3, TAG_ACCEPT = 4, TAG_PERMUTE = 5, TAG_PERMUTE_ACCEPT = 6 };

// Counter layout: c0 = move, c1 = tag<<24 | level<<16 | attempt, c2 = index,
// c3 = replica stream; key = seed (low word, high word).
void rng_block(uint64_t seed, uint32_t stream, uint32_t move, int tag, int level,
               uint32_t attempt, uint32_t index, uint32_t out[4]) {
  uint32_t c[4] = {move, ((uint32_t)tag << 24) | (((uint32_t)level & 0xffu) << 16) | (attempt & 0xffffu),
                   index, stream};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox(c, k, out);
}

double rng_uniform(uint64_t seed, uint32_t stream, uint32_t move, int tag, int level) {
  uint32_t w[4];
  rng_block(seed, stream, move, tag, level, 0, 0, w);
  return uniform53(w[0], w[1]);
}

// Polar Box-Muller (Random.h:156-180) on the pair's own sub-stream.
int gaussian_pair(uint64_t seed, uint32_t stream, uint32_t move, int level, uint32_t pair,
                  double g[2]) {
  for (uint32_t a = 0; a < 0x10000u; ++a) {
    uint32_t w[4];
    rng_block(seed, stream, move, TAG_GAUSS, level, a, pair, w);
    double V1 = 2.0 * uniform53(w[0], w[1]) - 1.0;
    double V2 = 2.0 * uniform53(w[2], w[3]) - 1.0;
    double S = V1 * V1 + V2 * V2;
    if (S < 1.0 && S > 0.0) {
      double temp = sqrt((-2.0 * log(S)) / S);
      g[0] = V1 * temp;
      g[1] = V2 * temp;
      return (int)a + 1;
    }
  }
  g[0] = g[1] = 0.0;
  return 0x10000;
}

// ------------------------------------------------------------------------
// Grids (Common/Splines/Grid.h)
// ------------------------------------------------------------------------
struct Grid {
  int type = PIMC_GRID_LINEAR;
  int n = 0;
  double start = 0, end = 0;
  std::vector<double> x;
  double delta = 0, deltainv = 0;  // LinearGrid
  double r0 = 0, spacing = 0;      // LogGrid (Z = 1)

  // LinearGrid::Init (Grid.h:87-96)
  void init_linear(double s, double e, int np) {
    type = PIMC_GRID_LINEAR;
    start = s; end = e; n = np;
    x.resize(n);
    delta = (end - start) / (double)(n - 1);
    deltainv = 1.0 / delta;
    for (int i = 0; i < n; i++) x[i] = start + (double)i * delta;
  }
  // The .dm reader's LOG grid (DavidPAClass.cc:714-719) -> LogGrid::Init (Grid.h:537-547)
  void init_log_dm(double s, double e, int np) {
    type = PIMC_GRID_LOG;
    double d = pow((e / s), 1.0 / (np - 1.0));
    n = np; r0 = s; spacing = d;
    start = r0;
    end = r0 * pow(spacing, (double)n - 1);
    x.resize(n);
    for (int i = 0; i < n; i++) x[i] = r0 * pow(spacing, (double)i);
  }
  static int to_int(double t, int n) {  // (int) cast of a value clamped to [-2, n+1]
    if (!(t >= -2.0)) t = -2.0;
    if (t > (double)(n + 1)) t = (double)(n + 1);
    return (int)t;
  }
  int reverse_map_literal(double r) const {
    if (type == PIMC_GRID_LINEAR) return to_int(nearbyint((r - start) * deltainv - 0.5), n);
    return to_int(floor(log(r / r0) / log(spacing)), n);
  }
  int reverse_map(double r) const {
    int i = reverse_map_literal(r);
    if (type == PIMC_GRID_LOG) {
      if (i >= 1 && i <= n - 2 && r < x[i]) i -= 1;
      else if (i >= -1 && i <= n - 3 && r >= x[i + 1]) i += 1;
    }
    return i;
  }
};

// ------------------------------------------------------------------------
// MultiCubicSpline (CubicSpline.h:470-561, CubicSpline.cc:77-128)
// ------------------------------------------------------------------------
struct MultiCubicSpline {
  const Grid* grid = nullptr;
  int ng = 0, ns = 0;
  std::vector<double> y, d2y;  // [ng][ns]
  std::vector<double> start_deriv, end_deriv;

  void init(const Grid* g, const double* ys, int nsplines, const double* sd, const double* ed) {
    grid = g; ng = g->n; ns = nsplines;
    y.assign(ys, ys + (size_t)ng * ns);
    d2y.assign((size_t)ng * ns, 0.0);
    start_deriv.assign(sd, sd + ns);
    end_deriv.assign(ed, ed + ns);
    for (int i = 0; i < ns; i++) update(i);
  }
  double& Y(int j, int i) { return y[(size_t)j * ns + i]; }
  double& D(int j, int i) { return d2y[(size_t)j * ns + i]; }

  void update(int i) {
    const std::vector<double>& x = grid->x;
    int N = ng;
    std::vector<double> U(N);
    if (start_deriv[i] > 0.99e30) {
      D(0, i) = 0.0;
      U[0] = 0.0;
    } else {
      D(0, i) = -0.5;
      U[0] = (3.0 / (x[1] - x[0])) * ((Y(1, i) - Y(0, i)) / (x[1] - x[0]) - start_deriv[i]);
    }
    D(N - 1, i) = 0.0;
    for (int j = 1; j < N - 1; j++) {
      double sig = (x[j] - x[j - 1]) / (x[j + 1] - x[j - 1]);
      double p = sig * D(j - 1, i) + 2.0;
      D(j, i) = (sig - 1.0) / p;
      U[j] = (6.0 * ((Y(j + 1, i) - Y(j, i)) / (x[j + 1] - x[j]) -
                     (Y(j, i) - Y(j - 1, i)) / (x[j] - x[j - 1])) /
                  (x[j + 1] - x[j - 1]) -
              sig * U[j - 1]) /
             p;
    }
    double Un, Qn;
    if (end_deriv[i] > 0.99e30) {
      Un = 0.0;
      Qn = 0.0;
    } else {
      Qn = 0.5;
      Un = (3.0 / (x[N - 1] - x[N - 2])) * (end_deriv[i] - (Y(N - 1, i) - Y(N - 2, i)) / (x[N - 1] - x[N - 2]));
    }
    D(N - 1, i) = (Un - Qn * U[N - 2]) / (Qn * D(N - 2, i) + 1.0);
    for (int k = N - 2; k >= 0; k--) D(k, i) = D(k, i) * D(k + 1, i) + U[k];
  }

  void locate(double xv, int& low, int& hi) const {
    hi = grid->reverse_map(xv) + 1;
    low = hi - 1;
    if (low < 0) { low = 0; hi = 1; }
    if (hi > (grid->n - 1)) { hi = grid->n - 1; low = hi - 1; }
  }
  double eval(int i, double xv) const {
    int low, hi;
    locate(xv, low, hi);
    const std::vector<double>& X = grid->x;
    double h = X[hi] - X[low];
    double hinv = 1.0 / h;
    double a = (X[hi] - xv) * hinv;
    double b = (xv - X[low]) * hinv;
    double sixinv = 0.1666666666666666666;
    return (a * y[(size_t)low * ns + i] + b * y[(size_t)hi * ns + i] +
            ((a * a * a - a) * d2y[(size_t)low * ns + i] + (b * b * b - b) * d2y[(size_t)hi * ns + i]) *
                (h * h * sixinv));
  }
  void eval_all(double xv, double* out) const {
    int low, hi;
    locate(xv, low, hi);
    const std::vector<double>& X = grid->x;
    double h = X[hi] - X[low];
    double hinv = 1.0 / h;
    double a = (X[hi] - xv) * hinv;
    double b = (xv - X[low]) * hinv;
    double a3minusa = a * a * a - a;
    double b3minusb = b * b * b - b;
    double h2over6 = h * h * 0.1666666666666666666;
    for (int i = 0; i < ns; i++)
      out[i] = a * y[(size_t)low * ns + i] + b * y[(size_t)hi * ns + i] +
               (a3minusa * d2y[(size_t)low * ns + i] + b3minusb * d2y[(size_t)hi * ns + i]) * h2over6;
  }
};

// ------------------------------------------------------------------------
// PairActionFitClass (PAFitBase.h:29-138): the part of the interface the path uses
// ------------------------------------------------------------------------
struct PairFit {
  virtual ~PairFit() {}
  virtual double U(double q, double z, double s2, int level) const = 0;
  // beta derivative (PairActionFitClass::dU); valid where has_dU()
  virtual double dU(double q, double z, double s2, int level) const = 0;
  virtual bool has_dU() const = 0;
  // the pair potential the fit was made from (PairActionFitClass::V); NaN where not stored
  virtual double V(double r) const { (void)r; return std::numeric_limits<double>::quiet_NaN(); }
  virtual bool IsLongRange() const = 0;
  // bisection levels the fit covers
  virtual int num_levels() const = 0;
};

// ------------------------------------------------------------------------
// BicubicSpline (BicubicSpline.h:24-30, 266-300, 356-413; BicubicSpline.cc:20-154)
// ------------------------------------------------------------------------
struct BCpoint {
  double z, dzdx, dzdy, d2zdxdy;
};

inline double p1(double t) { return ((t - 1.0) * (t - 1.0) * (1.0 + 2.0 * t)); }
inline double p2(double t) { return (t * t * (3.0 - 2.0 * t)); }
inline double q1(double t) { return (t * (t - 1.0) * (t - 1.0)); }
inline double q2(double t) { return (t * t * (t - 1.0)); }

struct BicubicSpline {
  const Grid* Xgrid = nullptr;
  const Grid* Ygrid = nullptr;
  int Nx = 0, Ny = 0;
  std::vector<BCpoint> Fv;
  BCpoint& F(int i, int j) { return Fv[(size_t)i * Ny + j]; }
  const BCpoint& F(int i, int j) const { return Fv[(size_t)i * Ny + j]; }

  void init(const Grid* xg, const Grid* yg, const double* f) {
    Xgrid = xg;
    Ygrid = yg;
    Nx = xg->n;
    Ny = yg->n;
    Fv.assign((size_t)Nx * Ny, BCpoint{0, 0, 0, 0});
    for (int i = 0; i < Nx; i++)
      for (int j = 0; j < Ny; j++) F(i, j).z = f[(size_t)i * Ny + j];
    for (int i = 0; i < Ny; i++) XUpdate(i);
    for (int i = 0; i < Nx; i++) YUpdate(i);
    BiUpdate();
  }
  void XUpdate(int iy) {
    const std::vector<double>& x = Xgrid->x;
    std::vector<double> mu(Nx);
    F(0, iy).dzdx = 1.5 * (F(1, iy).z - F(0, iy).z) / (x[1] - x[0]);
    F(Nx - 1, iy).dzdx = 1.5 * (F(Nx - 1, iy).z - F(Nx - 2, iy).z) / (x[Nx - 1] - x[Nx - 2]);
    mu[0] = 0.5;
    for (int j = 1; j < (Nx - 1); j++) {
      double lambda = 0.5 * (x[j + 1] - x[j]) / (x[j + 1] - x[j - 1]);
      mu[j] = 0.5 - lambda;
      F(j, iy).dzdx = 3.0 * (lambda * (F(j, iy).z - F(j - 1, iy).z) / (x[j] - x[j - 1]) +
                             mu[j] * (F(j + 1, iy).z - F(j, iy).z) / (x[j + 1] - x[j]));
      double cj = 1.0 - lambda * mu[j - 1];
      F(j, iy).dzdx -= lambda * F(j - 1, iy).dzdx;
      mu[j] /= cj;
      F(j, iy).dzdx /= cj;
    }
    int j = Nx - 1;
    double lambda = 0.5;
    mu[j] = 0.5 - lambda;
    F(j, iy).dzdx = 3.0 * (lambda * (F(j, iy).z - F(j - 1, iy).z) / (x[j] - x[j - 1]));
    double cj = 1.0 - lambda * mu[j - 1];
    F(j, iy).dzdx -= lambda * F(j - 1, iy).dzdx;
    mu[j] /= cj;
    F(j, iy).dzdx /= cj;
    for (j = Nx - 2; j >= 0; j--) F(j, iy).dzdx -= mu[j] * F(j + 1, iy).dzdx;
  }
  void YUpdate(int ix) {
    const std::vector<double>& y = Ygrid->x;
    std::vector<double> mu(Ny);
    F(ix, 0).dzdy = 1.5 * (F(ix, 1).z - F(ix, 0).z) / (y[1] - y[0]);
    F(ix, Ny - 1).dzdy = 1.5 * (F(ix, Ny - 1).z - F(ix, Ny - 2).z) / (y[Ny - 1] - y[Ny - 2]);
    mu[0] = 0.5;
    for (int j = 1; j < (Ny - 1); j++) {
      double lambda = 0.5 * (y[j + 1] - y[j]) / (y[j + 1] - y[j - 1]);
      mu[j] = 0.5 - lambda;
      F(ix, j).dzdy = 3.0 * (lambda * (F(ix, j).z - F(ix, j - 1).z) / (y[j] - y[j - 1]) +
                             mu[j] * (F(ix, j + 1).z - F(ix, j).z) / (y[j + 1] - y[j]));
      double cj = 1.0 - lambda * mu[j - 1];
      F(ix, j).dzdy -= lambda * F(ix, j - 1).dzdy;
      mu[j] /= cj;
      F(ix, j).dzdy /= cj;
    }
    int j = Ny - 1;
    double lambda = 0.5;
    mu[j] = 0.5 - lambda;
    F(ix, j).dzdy = 3.0 * (lambda * (F(ix, j).z - F(ix, j - 1).z) / (y[j] - y[j - 1]));
    double cj = 1.0 - lambda * mu[j - 1];
    F(ix, j).dzdy -= lambda * F(ix, j - 1).dzdy;
    mu[j] /= cj;
    F(ix, j).dzdy /= cj;
    for (j = Ny - 2; j >= 0; j--) F(ix, j).dzdy -= mu[j] * F(ix, j + 1).dzdy;
  }
  void BiUpdate() {
    const std::vector<double>& x = Xgrid->x;
    std::vector<double> mu(Nx);
    for (int iy = 0; iy < Ny; iy++) {
      F(0, iy).d2zdxdy = 1.5 * (F(1, iy).dzdy - F(0, iy).dzdy) / (x[1] - x[0]);
      F(Nx - 1, iy).d2zdxdy = 1.5 * (F(Nx - 1, iy).dzdy - F(Nx - 2, iy).dzdy) / (x[Nx - 1] - x[Nx - 2]);
      mu[0] = 0.5;
      for (int j = 1; j < (Nx - 1); j++) {
        double lambda = 0.5 * (x[j + 1] - x[j]) / (x[j + 1] - x[j - 1]);
        mu[j] = 0.5 - lambda;
        F(j, iy).d2zdxdy = 3.0 * (lambda * (F(j, iy).dzdy - F(j - 1, iy).dzdy) / (x[j] - x[j - 1]) +
                                  mu[j] * (F(j + 1, iy).dzdy - F(j, iy).dzdy) / (x[j + 1] - x[j]));
        double cj = 1.0 - lambda * mu[j - 1];
        F(j, iy).d2zdxdy -= lambda * F(j - 1, iy).d2zdxdy;
        mu[j] /= cj;
        F(j, iy).d2zdxdy /= cj;
      }
      int j = Nx - 1;
      double lambda = 0.5;
      mu[j] = 0.5 - lambda;
      F(j, iy).d2zdxdy = 3.0 * (lambda * (F(j, iy).dzdy - F(j - 1, iy).dzdy) / (x[j] - x[j - 1]));
      double cj = 1.0 - lambda * mu[j - 1];
      F(j, iy).d2zdxdy -= lambda * F(j - 1, iy).d2zdxdy;
      mu[j] /= cj;
      F(j, iy).d2zdxdy /= cj;
      for (j = Nx - 2; j >= 0; j--) F(j, iy).d2zdxdy -= mu[j] * F(j + 1, iy).d2zdxdy;
    }
  }
  double operator()(double x, double y) const {
    int ix = Xgrid->reverse_map(x);
    int iy = Ygrid->reverse_map(y);
    ix = std::max(0, ix);
    ix = std::min(ix, Nx - 2);
    iy = std::max(0, iy);
    iy = std::min(iy, Ny - 2);
    double h = Xgrid->x[ix + 1] - Xgrid->x[ix];
    double k = Ygrid->x[iy + 1] - Ygrid->x[iy];
    double u = (x - Xgrid->x[ix]) / h;
    double v = (y - Ygrid->x[iy]) / k;
    double a0 = p1(u), a1 = p2(u), a2 = h * q1(u), a3 = h * q2(u);
    double b0 = p1(v), b1 = p2(v), b2 = k * q1(v), b3 = k * q2(v);
    const BCpoint &A = F(ix, iy), &B = F(ix, iy + 1), &Cc = F(ix + 1, iy), &D = F(ix + 1, iy + 1);
    double val = a0 * (A.z * b0 + B.z * b1 + A.dzdy * b2 + B.dzdy * b3) +
                 a1 * (Cc.z * b0 + D.z * b1 + Cc.dzdy * b2 + D.dzdy * b3) +
                 a2 * (A.dzdx * b0 + B.dzdx * b1 + A.d2zdxdy * b2 + B.d2zdxdy * b3) +
                 a3 * (Cc.dzdx * b0 + D.dzdx * b1 + Cc.d2zdxdy * b2 + D.d2zdxdy * b3);
    return val;
  }
};

// ------------------------------------------------------------------------
// PAcoulombBCFitClass::U (PAcoulombBCFit.cc:159-176)
// ------------------------------------------------------------------------
struct CoulombBCFit : PairFit {
  Grid qgrid, tgrid;
  std::vector<BicubicSpline> Usplines, dUsplines;
  double SmallestBeta = 0, Z1Z2 = 0;
  // PairActionFitClass::V (PAFitBase.h:82): the Coulomb potential
  double V(double r) const override { return Z1Z2 / r; }
  // PAcoulombBCFitClass::dU (PAcoulombBCFit.cc:178-192): beyond the grid the Coulomb potential
  double dU(double q, double z, double s2, int level) const override {
    (void)z;
    if (q <= (qgrid.end * 1.0000001)) {
      if (q == 0) return (dUsplines[level](0.0, 0.0));
      double t = sqrt(s2) / (2.0 * q);
      return (dUsplines[level](q, t));
    }
    return Z1Z2 / q;
  }
  bool has_dU() const override { return !dUsplines.empty(); }
  double U(double q, double z, double s2, int level) const override {
    if (q <= (qgrid.end * 1.0000001)) {
      if (q == 0) return (Usplines[level](0.0, 0.0));
      double t = sqrt(s2) / (2.0 * q);
      return (Usplines[level](q, t));
    }
    double beta = SmallestBeta;
    for (int i = 0; i < level; i++) beta *= 2.0;
    return (beta * (Z1Z2 / q));   // CoulombPot::V
  }
  bool IsLongRange() const override { return true; }
  int num_levels() const override { return (int)Usplines.size(); }
};

// ------------------------------------------------------------------------
// QuinticSpline evaluation (QuinticSpline.h:172-218) on a LinearGrid
// ------------------------------------------------------------------------
struct QuinticTable {
  bool set = false;
  Grid grid;
  int levels = 0;
  std::vector<double> c;   // [levels][n][6] = Y, B, C, D, E, F
  double eval(int level, double x) const {
    int J = grid.reverse_map(x) + 1;
    int I = J - 1;
    if (I < 0) { I = 0; J = 1; }
    if (J > (grid.n - 1)) { J = grid.n - 1; I = J - 1; }
    double P = (x - grid.x[I]);
    const double* k = &c[((size_t)level * grid.n + I) * 6];
    return ((((k[5] * P + k[4]) * P + k[3]) * P + k[2]) * P + k[1]) * P + k[0];
  }
};

// ------------------------------------------------------------------------
// DavidPAClass (DavidPAClass.cc)
// ------------------------------------------------------------------------
struct DavidFit : PairFit {
  Grid grid;
  std::vector<MultiCubicSpline> ukj;  // per tau level
  std::vector<MultiCubicSpline> dukj; // beta derivatives, per tau level (empty when not given)
  int num_tau = 0, ncols = 0, num_terms = 0, taupos = 0;
  double smallest_tau = 0;
  mutable std::vector<double> temp, dtemp;   // TempukjArray, TempdukjArray

  // calcUsqzFast (DavidPAClass.cc:468-511)
  double calc_usqz_fast(double s, double q, double z, int level) const {
    level = level + taupos;
    const MultiCubicSpline& sp = ukj[level];
    double rmin = sp.grid->start;
    double U = 0.0;
    double r = q + 0.5 * z;
    double rprime = q - 0.5 * z;
    if (r > sp.grid->end) r = sp.grid->end;
    if (rprime > sp.grid->end) rprime = sp.grid->end;
    if ((rprime < rmin) || (r < rmin)) {
      U = sp.eval(1, rmin);
      return U;
    }
    U += 0.5 * (sp.eval(1, r) + sp.eval(1, rprime));
    if (s > 0.0 && q < sp.grid->end) {
      double zsquared = z * z;
      double ssquared = s * s;
      double ssquaredinverse = 1.0 / ssquared;
      double Sto2k = ssquared;
      sp.eval_all(q, temp.data());
      for (int k = 1; k <= num_terms; k++) {
        double Zto2j = 1;
        double currS = Sto2k;
        for (int j = 0; j <= k; j++) {
          double Ucof = temp[k * (k + 1) / 2 + j + 1];
          U += (Ucof)*Zto2j * currS;
          Zto2j *= zsquared;
          currS = currS * ssquaredinverse;
        }
        Sto2k = Sto2k * ssquared;
      }
    }
    return U;
  }
  // DavidPAClass::U (DavidPAClass.cc:92-102)
  double U(double q, double z, double s2, int level) const override {
    double s = sqrt(s2);
    return calc_usqz_fast(s, q, z, level);
  }
  // calcUsqz (DavidPAClass.cc:291-463): U, dU (dukj plus the potential the .dm
  // file leaves out of the diagonal) and V, with r, r' clamped into the grid.
  void calc_usqz(double s, double q, double z, int level, double& U, double& dUo, double& Vo) const {
    level = level + taupos;
    const MultiCubicSpline& sp = ukj[level];
    const MultiCubicSpline& dsp = dukj[level];
    double rmin = sp.grid->start;
    U = 0.0;
    dUo = 0.0;
    double r = q + 0.5 * z;
    double rprime = q - 0.5 * z;
    if (r > sp.grid->end) r = sp.grid->end;
    if (rprime > sp.grid->end) rprime = sp.grid->end;
    if (rprime < rmin) rprime = rmin;
    if (r < rmin) r = rmin;
    Vo = 0.5 * (sp.eval(0, r) + sp.eval(0, rprime));
    U += 0.5 * (sp.eval(1, r) + sp.eval(1, rprime));
    dUo += 0.5 * (dsp.eval(1, r) + dsp.eval(1, rprime));
    dUo += Vo;
    if (s > 0.0 && q < sp.grid->end) {
      double zsquared = z * z;
      double ssquared = s * s;
      double ssquaredinverse = 1.0 / ssquared;
      double Sto2k = ssquared;
      sp.eval_all(q, temp.data());
      dsp.eval_all(q, dtemp.data());
      for (int k = 1; k <= num_terms; k++) {
        double Zto2j = 1;
        double currS = Sto2k;
        for (int j = 0; j <= k; j++) {
          double Ucof = temp[k * (k + 1) / 2 + j + 1];
          double dUcof = dtemp[k * (k + 1) / 2 + j + 1];
          U += (Ucof)*Zto2j * currS;
          dUo += (dUcof)*Zto2j * currS;
          Zto2j *= zsquared;
          currS = currS * ssquaredinverse;
        }
        Sto2k = Sto2k * ssquared;
      }
    }
  }
  // DavidPAClass::dU (DavidPAClass.cc:115-126)
  double dU(double q, double z, double s2, int level) const override {
    double s = sqrt(s2);
    double u, du, v;
    calc_usqz(s, q, z, level, u, du, v);
    return du;
  }
  bool has_dU() const override { return !dukj.empty(); }
  // DavidPAClass::V (DavidPAClass.cc:104-113): column 0 of the first level, r clamped to the grid
  double V(double r) const override {
    if (r > ukj[0].grid->end) r = ukj[0].grid->end;
    if (r < ukj[0].grid->start) r = ukj[0].grid->start;
    return ukj[0].eval(0, r);
  }
  bool IsLongRange() const override { return false; }   // DavidPAClass.cc:127-130
  int num_levels() const override { return num_tau - taupos; }
};

// ------------------------------------------------------------------------
// Replica: PathClass state + action configuration
// ------------------------------------------------------------------------
struct LongRangeTables {
  bool set = false, dset = false;
  int num_levels = 0;
  std::vector<double> ulong_k;  // [levels][nk]
  std::vector<double> ulong_r0, ushort_k0, ulong_k0;
  std::vector<double> dulong_k, dulong_r0, dushort_k0, dulong_k0;   // beta derivatives
};

struct Replica {
  pimc_bisection_stats acc{};   // counters accumulated until pimc_bisection_stats_take
  int N = 0, M = 0, nslices = 0, ns = 0, num_images = 0;
  double box[3], boxinv[3], isperiodic[3];
  double tau = 0;
  std::vector<int> species_of, first, count;
  std::vector<double> lambda;
  std::vector<int> pair_fit;     // [ns*ns]
  std::vector<double> path[2];   // [slice][ptcl][3]
  std::vector<int> perm[2];
  bool long_range = false, use_long_range = false, use_background = false;
  double kcutoff = 0;
  double kbox[3];
  int maxk[3];
  std::vector<double> kvecs;     // [nk][3]
  std::vector<int> kidx;         // [nk][3]
  int nk = 0;
  std::vector<cplx> rho[2];      // [slice][species][k]
  std::vector<LongRangeTables> lr;  // [ns*ns]
  std::vector<QuinticTable> ulong;  // [ns*ns] PairActionFitClass::Ulong of the pair
  std::vector<QuinticTable> dulong; // [ns*ns] dUlong
  // [ns*ns] PairActionFitClass Vlong (a one-level QuinticTable), Vlong_k, Vlong_r0,
  // Vshort_k0, Vlong_k0 (PAFitBase.h:39-53)
  struct VlongTables {
    bool set = false;
    QuinticTable spline;
    std::vector<double> vk;
    double r0 = 0, short_k0 = 0, long_k0 = 0;
  };
  std::vector<VlongTables> vlong;
  uint32_t stream = 0;
  int join = 0;
  // PermuteType TABLE for moves of `permute_species` (TablePermuteStageClass::Read)
  bool permute_on = false;
  int permute_species = 0;
  double permute_gamma[4] = {0, 0, 0, 0};
  double permute_epsilon = 0;

  double* P(int mode, int slice, int ptcl) { return &path[mode][((size_t)slice * N + ptcl) * 3]; }
  cplx& R(int mode, int slice, int s, int k) { return rho[mode][((size_t)slice * ns + s) * nk + k]; }

  // PathClass::PutInBox (PathClass.h:897-921)
  void put_in_box(double v[3]) const {
    for (int i = 0; i < 3; i++) {
      double n = -floor(v[i] * boxinv[i] + 0.5);
      v[i] += n * isperiodic[i] * box[i];
    }
  }
  // PathClass::PutInBoxFast (PathClass.cc:1389-1407)
  void put_in_box_fast(double v[3]) const {
    for (int d = 0; d < 3; d++) {
      double half = box[d] / 2;
      if (v[d] > half) v[d] -= box[d];
      if (v[d] < -half) v[d] += box[d];
    }
  }
  // PathClass::Velocity (PathClass.h:831-865)
  void velocity(int mode, int sA, int sB, int p, double vel[3]) {
    const double* a = P(mode, sA, p);
    const double* b = P(mode, sB, p);
    for (int d = 0; d < 3; d++) vel[d] = b[d] - a[d];
    put_in_box(vel);
  }
  // PathClass::DistDisp(sliceA, sliceB, ...) (PathClass.h:696-746)
  void dist_disp(int mode, int sA, int sB, int p1, int p2, double& distA, double& distB,
                 double dispA[3], double dispB[3], double nimg[3], double mimg[3]) {
    const double* a1 = P(mode, sA, p1);
    const double* a2 = P(mode, sA, p2);
    const double* b1 = P(mode, sB, p1);
    const double* b2 = P(mode, sB, p2);
    for (int d = 0; d < 3; d++) {
      dispA[d] = a2[d] - a1[d];
      dispB[d] = b2[d] - b1[d];
    }
    double n[3], m[3];
    for (int d = 0; d < 3; d++) n[d] = -nearbyint(dispA[d] * boxinv[d]);
    for (int d = 0; d < 3; d++) dispA[d] += n[d] * isperiodic[d] * box[d];
    for (int d = 0; d < 3; d++) m[d] = nearbyint((dispA[d] - dispB[d]) * boxinv[d]);
    for (int d = 0; d < 3; d++) dispB[d] += m[d] * isperiodic[d] * box[d];
    distA = sqrt(dispA[0] * dispA[0] + dispA[1] * dispA[1] + dispA[2] * dispA[2]);
    distB = sqrt(dispB[0] * dispB[0] + dispB[1] * dispB[1] + dispB[2] * dispB[2]);
    if (nimg) for (int d = 0; d < 3; d++) nimg[d] = n[d];
    if (mimg) for (int d = 0; d < 3; d++) mimg[d] = m[d];
  }

  // PathClass::SetupkVecs3D + Include (PathClass.cc:281-292,534-586)
  void setup_kvecs() {
    for (int i = 0; i < 3; i++) kbox[i] = 2.0 * M_PI / box[i];
    for (int i = 0; i < 3; i++) maxk[i] = (int)ceil(1.1 * kcutoff / kbox[i]);
    kvecs.clear();
    kidx.clear();
    for (int ix = -maxk[0]; ix <= maxk[0]; ix++) {
      double k0 = ix * kbox[0];
      for (int iy = -maxk[1]; iy <= maxk[1]; iy++) {
        double k1 = iy * kbox[1];
        for (int iz = -maxk[2]; iz <= maxk[2]; iz++) {
          double k2 = iz * kbox[2];
          double kk = k0 * k0 + k1 * k1 + k2 * k2;
          bool include = (k0 > 0.0) || ((k0 == 0.0) && (k1 > 0.0)) ||
                         ((k0 == 0.0) && (k1 == 0.0) && (k2 > 0.0));
          if ((kk < kcutoff * kcutoff) && include) {
            kvecs.push_back(k0); kvecs.push_back(k1); kvecs.push_back(k2);
            kidx.push_back(ix + maxk[0]); kidx.push_back(iy + maxk[1]); kidx.push_back(iz + maxk[2]);
          }
        }
      }
    }
    nk = (int)(kvecs.size() / 3);
  }

  // The C_d(n) = e^{i n phi_d} tables of CalcRho_ks_Fast / UpdateRho_ks.
  void phase_tables(const double* r, std::vector<cplx> C[3]) const {
    for (int dim = 0; dim < 3; dim++) {
      C[dim].assign(2 * maxk[dim] + 1, cplx(0.0, 0.0));
      double phi = r[dim] * kbox[dim];
      cplx tempC(cos(phi), sin(phi));
      C[dim][maxk[dim]] = 1.0;
      for (int n = 1; n <= maxk[dim]; n++) {
        C[dim][maxk[dim] + n] = tempC * C[dim][maxk[dim] + n - 1];
        C[dim][maxk[dim] - n] = std::conj(C[dim][maxk[dim] + n]);
      }
    }
  }
  // PathClass::CalcRho_ks_Fast (PathClass.cc:616-647)
  void calc_rho_fast(int mode, int slice, int species) {
    for (int ki = 0; ki < nk; ki++) R(mode, slice, species, ki) = 0.0;
    std::vector<cplx> C[3];
    for (int ptcl = first[species]; ptcl < first[species] + count[species]; ptcl++) {
      phase_tables(P(mode, slice, ptcl), C);
      for (int ki = 0; ki < nk; ki++) {
        const int* kI = &kidx[3 * ki];
        R(mode, slice, species, ki) += C[0][kI[0]] * C[1][kI[1]] * C[2][kI[2]];
      }
    }
  }
  // PathClass::UpdateRho_ks() (PathClass.cc:709-723)
  void update_rho_all() {
    for (int slice = 0; slice < nslices; slice++)
      for (int s = 0; s < ns; s++) calc_rho_fast(PIMC_OLDMODE, slice, s);
    rho[1] = rho[0];
  }
  // PathClass::UpdateRho_ks(slice1, slice2, changed, level) (PathClass.cc:649-707)
  void update_rho(int slice1, int slice2, const int* ptcls, int nptcls, int level) {
    int skip = 1 << level;
    for (int slice = slice1; slice <= slice2; slice += skip)
      for (int s = 0; s < ns; s++)
        for (int ki = 0; ki < nk; ki++) R(1, slice, s, ki) = R(0, slice, s, ki);
    std::vector<cplx> C[3];
    for (int pi = 0; pi < nptcls; pi++) {
      int ptcl = ptcls[pi];
      int species = species_of[ptcl];
      for (int slice = slice1; slice <= slice2; slice += skip) {
        phase_tables(P(1, slice, ptcl), C);
        for (int ki = 0; ki < nk; ki++) {
          const int* kI = &kidx[3 * ki];
          R(1, slice, species, ki) += C[0][kI[0]] * C[1][kI[1]] * C[2][kI[2]];
        }
        phase_tables(P(0, slice, ptcl), C);
        for (int ki = 0; ki < nk; ki++) {
          const int* kI = &kidx[3 * ki];
          R(1, slice, species, ki) -= C[0][kI[0]] * C[1][kI[1]] * C[2][kI[2]];
        }
      }
    }
  }
};

int make_grid_checked(int32_t type, double start, double end, int32_t npoints, Grid& g) {
  if (npoints < 2 || !(start < end)) return fail(PIMC_ERR_ARG, "grid: bad range or size");
  if (type == PIMC_GRID_LINEAR)
    g.init_linear(start, end, npoints);
  else if (type == PIMC_GRID_LOG && start > 0)
    g.init_log_dm(start, end, npoints);
  else
    return fail(PIMC_ERR_ARG, "grid: bad type");
  return PIMC_OK;
}

struct Ctx {
  std::vector<std::unique_ptr<PairFit>> fits;
  std::vector<std::unique_ptr<Replica>> reps;
};

Replica* get_rep(Ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->reps.size()) return nullptr;
  return c->reps[id].get();
}

// ------------------------------------------------------------------------
// Actions
// ------------------------------------------------------------------------

// ShortRangeClass::SingleAction (ShortRangeClass.cc:131-176)
double short_range_action(Ctx* c, Replica& R, int mode, int slice1, int slice2, const int* ptcls,
                          int nptcls, int level, int64_t* evals) {
  double TotalU = 0.0;
  std::vector<char> DoPtcl(R.N, 1);
  int skip = 1 << level;
  for (int i1 = 0; i1 < nptcls; i1++) {
    int ptcl1 = ptcls[i1];
    DoPtcl[ptcl1] = 0;
    int species1 = R.species_of[ptcl1];
    for (int ptcl2 = 0; ptcl2 < R.N; ptcl2++) {
      if (DoPtcl[ptcl2]) {
        int species2 = R.species_of[ptcl2];
        const PairFit& PA = *c->fits[R.pair_fit[species1 * R.ns + species2]];
        const QuinticTable& Ulong = R.ulong[species1 * R.ns + species2];
        const bool subtract = PA.IsLongRange() && R.use_long_range;
        for (int slice = slice1; slice < slice2; slice += skip) {
          double r[3], rp[3], rmag, rpmag;
          R.dist_disp(mode, slice, slice + skip, ptcl1, ptcl2, rmag, rpmag, r, rp, nullptr, nullptr);
          double d0 = r[0] - rp[0], d1 = r[1] - rp[1], d2 = r[2] - rp[2];
          double s2 = d0 * d0 + d1 * d1 + d2 * d2;
          double q = 0.5 * (rmag + rpmag);
          double z = (rmag - rpmag);
          double U = PA.U(q, z, s2, level);
          if (subtract) U -= 0.5 * (Ulong.eval(level, rmag) + Ulong.eval(level, rpmag));
          TotalU += U;
          if (evals) ++*evals;
        }
      }
    }
  }
  return TotalU;
}

// KineticClass::SingleAction (KineticClass.cc:33-90)
double kinetic_action(Replica& R, int mode, int slice1, int slice2, const int* ptcls, int nptcls,
                      int level) {
  double TotalK = 0.0;
  int skip = 1 << level;
  double levelTau = R.tau * (1 << level);
  for (int pi = 0; pi < nptcls; pi++) {
    int ptcl = ptcls[pi];
    int species = R.species_of[ptcl];
    double lambda = R.lambda[species];
    if (lambda != 0) {
      double FourLambdaTauInv = 1.0 / (4.0 * lambda * levelTau);
      for (int slice = slice1; slice < slice2; slice += skip) {
        double vel[3];
        R.velocity(mode, slice, slice + skip, ptcl, vel);
        double GaussProd = 1.0;
        for (int dim = 0; dim < 3; dim++) {
          double GaussSum = 0.0;
          for (int image = -R.num_images; image <= R.num_images; image++) {
            double dist = vel[dim] + (double)image * R.box[dim];
            GaussSum += exp(-dist * dist * FourLambdaTauInv);
          }
          GaussProd *= GaussSum;
        }
        TotalK -= log(GaussProd);
      }
    }
  }
  return TotalK;
}

inline double mag2(const cplx& z) { return z.real() * z.real() + z.imag() * z.imag(); }

// LongRangeClass::SingleAction (LongRangeClass.cc:273-371), without its debug output
double long_range_action(Replica& R, int mode, int slice1, int slice2, const int* ptcls, int nptcls,
                         int level, int64_t* kevals) {
  if (R.nk == 0) return 0.0;
  if (mode == PIMC_NEWMODE) R.update_rho(slice1, slice2, ptcls, nptcls, level);
  int skip = 1 << level;
  double homo = 0.0, hetero = 0.0, background = 0.0, k0Homo = 0.0, k0Hetero = 0.0;
  for (int slice = slice1; slice <= slice2; slice += skip) {
    double factor = ((slice == slice1) || (slice == slice2)) ? 0.5 : 1.0;
    for (int species = 0; species < R.ns; species++) {
      const LongRangeTables& pa = R.lr[species * R.ns + species];
      if (pa.set) {
        for (int ki = 0; ki < R.nk; ki++) {
          double rhok2 = mag2(R.R(mode, slice, species, ki));
          homo += factor * 0.5 * 2.0 * rhok2 * pa.ulong_k[(size_t)level * R.nk + ki];
        }
        if (kevals) *kevals += R.nk;
        int N = R.count[species];
        homo -= factor * 0.5 * N * pa.ulong_r0[level];
        background -= factor * 0.5 * N * N * pa.ushort_k0[level];
        k0Homo += factor * 0.5 * N * N * pa.ulong_k0[level];
      }
    }
    for (int s1 = 0; s1 < R.ns; s1++)
      for (int s2 = s1 + 1; s2 < R.ns; s2++) {
        const LongRangeTables& pa = R.lr[s1 * R.ns + s2];
        if (pa.set) {
          for (int ki = 0; ki < R.nk; ki++) {
            const cplx& a = R.R(mode, slice, s1, ki);
            const cplx& b = R.R(mode, slice, s2, ki);
            double rhorho = a.real() * b.real() + a.imag() * b.imag();
            hetero += factor * 2.0 * rhorho * pa.ulong_k[(size_t)level * R.nk + ki];
          }
          if (kevals) *kevals += R.nk;
          int N1 = R.count[s1], N2 = R.count[s2];
          background -= factor * N1 * N2 * pa.ushort_k0[level];
          k0Hetero += factor * N1 * N2 * pa.ulong_k0[level];
        }
      }
  }
  double U = homo + hetero;
  if (R.use_background)
    U += background;
  else
    U += (k0Homo + k0Hetero);
  return U;
}

bool has_long_range(const Replica& R) {
  if (!R.use_long_range || R.nk == 0) return false;
  for (const auto& t : R.lr)
    if (t.set) return true;
  return false;
}

// PathClass::AcceptCopy / RejectCopy (PathClass.cc:797-899)
void accept_copy(Replica& R, int slice1, int slice2, const int* ptcls, int nptcls) {
  for (int pi = 0; pi < nptcls; pi++) {
    int ptcl = ptcls[pi];
    for (int slice = slice1; slice <= slice2; slice++) R.put_in_box_fast(R.P(1, slice, ptcl));
    for (int slice = slice1; slice <= slice2; slice++)
      memcpy(R.P(0, slice, ptcl), R.P(1, slice, ptcl), 3 * sizeof(double));
    R.perm[0][ptcl] = R.perm[1][ptcl];
  }
  if (R.nk > 0)
    for (int slice = slice1; slice <= slice2; slice++)
      for (int s = 0; s < R.ns; s++)
        for (int k = 0; k < R.nk; k++) R.R(0, slice, s, k) = R.R(1, slice, s, k);
}
void reject_copy(Replica& R, int slice1, int slice2, const int* ptcls, int nptcls) {
  for (int pi = 0; pi < nptcls; pi++) {
    int ptcl = ptcls[pi];
    for (int slice = slice1; slice <= slice2; slice++)
      memcpy(R.P(1, slice, ptcl), R.P(0, slice, ptcl), 3 * sizeof(double));
    R.perm[1][ptcl] = R.perm[0][ptcl];
  }
  if (R.nk > 0)
    for (int slice = slice1; slice <= slice2; slice++)
      for (int s = 0; s < R.ns; s++)
        for (int k = 0; k < R.nk; k++) R.R(1, slice, s, k) = R.R(0, slice, s, k);
}
// PathClass::MoveJoin (PathClass.cc:727-793), closed paths
void move_join(Replica& R, int oldJoin, int newJoin) {
  if (newJoin > oldJoin) {
    for (int t = oldJoin + 1; t <= newJoin; t++)
      for (int p = 0; p < R.N; p++) memcpy(R.P(0, t, p), R.P(1, t, R.perm[1][p]), 3 * sizeof(double));
    for (int t = oldJoin + 1; t <= newJoin; t++)
      for (int p = 0; p < R.N; p++) memcpy(R.P(1, t, p), R.P(0, t, p), 3 * sizeof(double));
  } else if (oldJoin > newJoin) {
    for (int t = newJoin + 1; t <= oldJoin; t++)
      for (int p = 0; p < R.N; p++) memcpy(R.P(0, t, R.perm[1][p]), R.P(1, t, p), 3 * sizeof(double));
    for (int t = newJoin + 1; t <= oldJoin; t++)
      for (int p = 0; p < R.N; p++) memcpy(R.P(1, t, p), R.P(0, t, p), 3 * sizeof(double));
  }
}

// ------------------------------------------------------------------------
// Bisection move (BisectionBlockClass::MakeMove with NoPermuteStage)
// ------------------------------------------------------------------------

// BisectionStageClass::Sample (BisectionStage.cc:128-289), no correlated sampling.
double bisection_sample(Replica& R, int slice1, int num_levels, int level, const int* ptcls,
                        int nptcls, uint64_t seed, uint32_t move) {
  int skip = 1 << (level + 1);
  double levelTau = 0.5 * R.tau * skip;
  int numImages = R.num_images;
  int slice2 = slice1 + (1 << num_levels);
  int nbeads = (slice2 - slice1) / skip;
  double logSampleProb = 0.0;
  double logOldSampleProb = 0.0;
  double gpair[2];
  int cached_pair = -1;
  for (int pi = 0; pi < nptcls; pi++) {
    int ptcl = ptcls[pi];
    double lambda = R.lambda[R.species_of[ptcl]];
    double sigma2 = (1.0 * lambda * levelTau);
    double sigma = sqrt(sigma2);
    double prefactorOfSampleProb = 0.0;
    int b = 0;
    for (int slice = slice1; slice < slice2; slice += skip, b++) {
      double rOld[3], rdiffOld[3], rbarOld[3], DeltaOld[3];
      memcpy(rOld, R.P(0, slice, ptcl), sizeof rOld);
      R.velocity(0, slice, slice + skip, ptcl, rdiffOld);
      for (int d = 0; d < 3; d++) rbarOld[d] = rOld[d] + 0.5 * rdiffOld[d];
      const double* rppOld = R.P(0, slice + (skip >> 1), ptcl);
      for (int d = 0; d < 3; d++) DeltaOld[d] = rppOld[d] - rbarOld[d];
      R.put_in_box(DeltaOld);

      double r[3], rdiff[3], rbar[3];
      memcpy(r, R.P(1, slice, ptcl), sizeof r);
      R.velocity(1, slice, slice + skip, ptcl, rdiff);
      for (int d = 0; d < 3; d++) rbar[d] = r[d] + 0.5 * rdiff[d];

      double Delta[3];
      for (int d = 0; d < 3; d++) {
        uint32_t gi = (uint32_t)(3 * (pi * nbeads + b) + d);
        uint32_t pair = gi >> 1;
        if ((int)pair != cached_pair) {
          gaussian_pair(seed, R.stream, move, level, pair, gpair);
          cached_pair = (int)pair;
        }
        Delta[d] = gpair[gi & 1] * sigma;  // LocalGaussian(sigma) = X1 * sigma
      }
      R.put_in_box(Delta);
      double GaussProd = 1.0, GaussProdOld = 1.0;
      for (int dim = 0; dim < 3; dim++) {
        double GaussSum = 0.0, GaussSumOld = 0.0;
        for (int image = -numImages; image <= numImages; image++) {
          double dist = Delta[dim] + (double)image * R.box[dim];
          double distOld = DeltaOld[dim] + (double)image * R.box[dim];
          GaussSum += exp(-0.5 * dist * dist / sigma2);
          GaussSumOld += exp(-0.5 * distOld * distOld / sigma2);
        }
        GaussProd *= GaussSum;
        GaussProdOld *= GaussSumOld;
      }
      logOldSampleProb += prefactorOfSampleProb + log(GaussProdOld);
      logSampleProb += prefactorOfSampleProb + log(GaussProd);
      double* rpp = R.P(1, slice + (skip >> 1), ptcl);
      for (int d = 0; d < 3; d++) rpp[d] = rbar[d] + Delta[d];
    }
  }
  return -logSampleProb + logOldSampleProb;
}

// StageClass::StageAction with [Kinetic, ShortRange, (LongRange at the lowest level)].
double stage_action(Ctx* c, Replica& R, int mode, int slice1, int slice2, const int* ptcls,
                    int nptcls, int level, bool with_long_range, pimc_bisection_stats* st) {
  double TotalAction = 0.0;
  TotalAction += kinetic_action(R, mode, slice1, slice2, ptcls, nptcls, level);
  int64_t evals = 0, kev = 0;
  TotalAction += short_range_action(c, R, mode, slice1, slice2, ptcls, nptcls, level, &evals);
  if (with_long_range) TotalAction += long_range_action(R, mode, slice1, slice2, ptcls, nptcls, level, &kev);
  if (st) {
    st->pair_evals += evals;
    st->k_evals += kev;
  }
  return TotalAction;
}

void table_move(Ctx* c, Replica& R, const pimc_bisection_desc& bd, uint32_t move, pimc_move_record* rec,
                pimc_bisection_stats* st);

void bisection_move(Ctx* c, Replica& R, const pimc_bisection_desc& bd, uint32_t move,
                    pimc_move_record* rec, pimc_bisection_stats* st) {
  if (R.permute_on && R.permute_species == bd.species) {
    table_move(c, R, bd, move, rec, st);
    return;
  }
  const int L = bd.num_levels;
  // BisectionBlockClass::ChooseTimeSlices (BisectionBlock.cc:363-374)
  int sliceSep = 1 << L;
  int numLeft = R.nslices - sliceSep;
  int slice1 = (int)floor(rng_uniform(bd.seed, R.stream, move, TAG_SLICE, 0) * numLeft);
  int slice2 = slice1 + sliceSep;
  move_join(R, R.join, slice2);
  R.join = slice2;
  // NoPermuteStageClass::ChooseParticle (NoPermuteClass.cc:61-68)
  int sp = bd.species;
  int ptcl = (int)floor(rng_uniform(bd.seed, R.stream, move, TAG_PTCL, 0) * R.count[sp]) + R.first[sp];
  int ptcls[1] = {ptcl};
  bool lr = has_long_range(R);
  if (rec) {
    memset(rec, 0, sizeof *rec);
    rec->slice1 = slice1;
    rec->ptcl = ptcl;
  }
  // MultiStageClass::MakeMove (MultiStage.cc:42-86) + LocalStageClass::Attempt (StageClass.cc:49-95)
  bool toAccept = true;
  double prevActionChange = 0.0;
  int stage = 0;
  for (int level = L - 1; level >= bd.lowest_level && toAccept; level--, stage++) {
    double logSampleRatio = bisection_sample(R, slice1, L, level, ptcls, 1, bd.seed, move);
    bool with_lr = lr && (level == bd.lowest_level);
    double oldAction = stage_action(c, R, PIMC_OLDMODE, slice1, slice2, ptcls, 1, level, with_lr, st);
    double newAction = stage_action(c, R, PIMC_NEWMODE, slice1, slice2, ptcls, 1, level, with_lr, st);
    double currActionChange = newAction - oldAction;
    double logAcceptProb = logSampleRatio - currActionChange + prevActionChange;
    double ran = rng_uniform(bd.seed, R.stream, move, TAG_ACCEPT, level);
    double logu = log(ran);
    toAccept = logAcceptProb >= logu;
    prevActionChange = currActionChange;
    if (st) st->stages++;
    if (rec) {
      rec->old_action[stage] = oldAction;
      rec->new_action[stage] = newAction;
      rec->log_sample_ratio[stage] = logSampleRatio;
      rec->log_u[stage] = logu;
      rec->num_stages = stage + 1;
    }
  }
  if (toAccept)
    accept_copy(R, slice1, slice2, ptcls, 1);
  else
    reject_copy(R, slice1, slice2, ptcls, 1);
  if (rec) rec->accepted = toAccept ? 1 : 0;
  if (st) {
    st->moves++;
    if (toAccept) st->accepted++;
  }
}

// ------------------------------------------------------------------------
// Bisection move with PermuteType TABLE (BisectionBlock.cc:51-64,126: the table permute
// stage first and last in the stage list)
// ------------------------------------------------------------------------

// One entry of PermuteTableClass::ConstructHTable / UpdateHTable (PermuteTableClass.cc:88-150)
// for species-relative particles i, j from copy `mode`: exp(-(|r_j(slice2) - r_i(slice1)|^2 -
// |r_i(slice2) - r_i(slice1)|^2) / (4 lambda beta)), stored as float, 0 below log(epsilon).
float htable_entry(Replica& R, int mode, int first, int i, int j, int slice1, int slice2, double logEps,
                   double fourLambdaBetaInv) {
  double disp_ii[3], disp_ij[3];
  for (int d = 0; d < 3; d++) disp_ii[d] = R.P(mode, slice2, i + first)[d] - R.P(mode, slice1, i + first)[d];
  R.put_in_box(disp_ii);
  const double dist_ii = disp_ii[0] * disp_ii[0] + disp_ii[1] * disp_ii[1] + disp_ii[2] * disp_ii[2];
  for (int d = 0; d < 3; d++) disp_ij[d] = R.P(mode, slice2, j + first)[d] - R.P(mode, slice1, i + first)[d];
  R.put_in_box(disp_ij);
  const double dist_ij = disp_ij[0] * disp_ij[0] + disp_ij[1] * disp_ij[1] + disp_ij[2] * disp_ij[2];
  const double toExp = (-dist_ij + dist_ii) * fourLambdaBetaInv;
  return toExp < logEps ? 0.0f : (float)exp(toExp);
}

struct CycleEntry {
  int len;
  int rep[4];
  float P, C;   // weight and running sum, in float as CycleClass holds them
};

// PermuteTableClass::ConstructBosonCycleTable (PermuteTableClass.cc:339-449) with no excluded
// particle and no z focus: the cycles in the reference's loop order, each weight
// Gamma[len-1] prod H along the cycle rounded to float, kept when above epsilon, with the
// running sum C formed in float.  Returns Norm (the last C); table may be null.
float boson_cycle_table(const float* H, int n, const double* Gamma, double epsilon, std::vector<CycleEntry>* table) {
  int num = 0;
  float last_c = 0.0f;
  auto add = [&](const CycleEntry& e) {
    if (table) table->push_back(e);
    last_c = e.C;
    num++;
  };
  for (int i = 0; i < n; i++) {
    CycleEntry t{};
    t.len = 1;
    t.rep[0] = i;
    t.P = (float)Gamma[0];
    t.C = t.P;
    if (num != 0) t.C += last_c;
    add(t);
    for (int j = i + 1; j < n; j++) {   // 2 and higher cycles
      t.rep[1] = j;
      const double hprod = H[(size_t)i * n + j];
      t.len = 2;
      t.P = (float)(Gamma[1] * hprod * H[(size_t)j * n + i]);
      t.C = t.P + last_c;
      if (t.P > epsilon) add(t);
      if (hprod * Gamma[2] * Gamma[3] > epsilon) {
        for (int k = i + 1; k < n; k++) {   // 3 and higher cycles
          if (k == j) continue;
          t.rep[2] = k;
          const double hprod2 = hprod * H[(size_t)j * n + k];
          t.len = 3;
          t.P = (float)(Gamma[2] * hprod2 * H[(size_t)k * n + i]);
          t.C = t.P + last_c;
          if (t.P > epsilon) add(t);
          if (hprod2 * Gamma[3] > epsilon) {
            for (int l = i + 1; l < n; l++) {
              if (l == j || l == k) continue;
              const double hprod3 = hprod2 * H[(size_t)k * n + l];
              t.rep[3] = l;
              t.len = 4;
              t.P = (float)(Gamma[3] * hprod3 * H[(size_t)l * n + i]);
              t.C = t.P + last_c;
              if (t.P > epsilon) add(t);
            }
          }
        }
      }
    }
  }
  return last_c;
}

// PermuteTableClass::FindEntry (PermuteTableClass.h:166-190): binary search of xi Norm in C.
int find_entry(const std::vector<CycleEntry>& table, double Norm, double xi) {
  xi *= Norm;
  int hi = (int)table.size() - 1, lo = 0;
  if (xi < table[0].C) return 0;
  while (hi - lo > 1) {
    const int attempt = (hi + lo) >> 1;
    if (xi < table[attempt].C)
      hi = attempt;
    else
      lo = attempt;
  }
  return hi;
}

// CycleClass::Apply (PermuteTableClass.cc:172-187) on the NEW copy at one slice.
void cycle_apply_new(Replica& R, int first, int slice, int len, const int* rep) {
  double tempPos[3];
  memcpy(tempPos, R.P(1, slice, rep[0] + first), sizeof tempPos);
  const int tempPtcl = R.perm[1][rep[0] + first];
  for (int i = 0; i < len - 1; i++) {
    memcpy(R.P(1, slice, rep[i] + first), R.P(1, slice, rep[i + 1] + first), 3 * sizeof(double));
    R.perm[1][rep[i] + first] = R.perm[1][rep[i + 1] + first];
  }
  memcpy(R.P(1, slice, rep[len - 1] + first), tempPos, sizeof tempPos);
  R.perm[1][rep[len - 1] + first] = tempPtcl;
}

void table_move(Ctx* c, Replica& R, const pimc_bisection_desc& bd, uint32_t move, pimc_move_record* rec,
                pimc_bisection_stats* st) {
  const int L = bd.num_levels;
  // BisectionBlockClass::ChooseTimeSlices, MoveJoin (BisectionBlock.cc:363-374, 424)
  const int sliceSep = 1 << L;
  const int numLeft = R.nslices - sliceSep;
  const int slice1 = (int)floor(rng_uniform(bd.seed, R.stream, move, TAG_SLICE, 0) * numLeft);
  const int slice2 = slice1 + sliceSep;
  move_join(R, R.join, slice2);
  R.join = slice2;
  const int sp = bd.species, first = R.first[sp], n = R.count[sp];
  const double* Gamma = R.permute_gamma;
  const double epsilon = R.permute_epsilon;
  // TablePermuteStageClass::InitBlock: Forw->ConstructHTable (TablePermuteStage.cc:127-140)
  const double logEps = log(epsilon);
  const double beta = R.tau * (double)(slice2 - slice1);
  const double fourLambdaBetaInv = 1.0 / (4.0 * R.lambda[sp] * beta);
  std::vector<float> H((size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      H[(size_t)i * n + j] = htable_entry(R, 1, first, i, j, slice1, slice2, logEps, fourLambdaBetaInv);
  // first Attempt: the cycle table and AttemptPermutation (TablePermuteStage.cc:204-230,
  // PermuteTableClass.cc:189-203)
  std::vector<CycleEntry> table;
  const double Norm = boson_cycle_table(H.data(), n, Gamma, epsilon, &table);
  const double NormInv = 1.0 / Norm;
  const double xi = rng_uniform(bd.seed, R.stream, move, TAG_PERMUTE, 0);
  const CycleEntry cyc = table[find_entry(table, Norm, xi)];
  cycle_apply_new(R, first, slice2, cyc.len, cyc.rep);
  const double forwT = cyc.P * NormInv;
  const int len = cyc.len;
  int ptcls[4];
  for (int i = 0; i < len; i++) ptcls[i] = cyc.rep[i] + first;
  double prevActionChange = -log(cyc.P / Gamma[len - 1]);
  if (rec) {
    memset(rec, 0, sizeof *rec);
    rec->slice1 = slice1;
    rec->ptcl = ptcls[0];
    rec->cycle_length = len;
    for (int i = 0; i < len; i++) rec->cycle[i] = ptcls[i];
    rec->num_cycles = (int)table.size();
    rec->forward_t = forwT;
  }
  // the bisection stages, moving every particle of the cycle (LocalStageClass::Attempt)
  const bool lr = has_long_range(R);
  bool toAccept = true;
  int stage = 0;
  for (int level = L - 1; level >= bd.lowest_level && toAccept; level--, stage++) {
    const double logSampleRatio = bisection_sample(R, slice1, L, level, ptcls, len, bd.seed, move);
    const bool with_lr = lr && (level == bd.lowest_level);
    const double oldAction = stage_action(c, R, PIMC_OLDMODE, slice1, slice2, ptcls, len, level, with_lr, st);
    const double newAction = stage_action(c, R, PIMC_NEWMODE, slice1, slice2, ptcls, len, level, with_lr, st);
    const double currActionChange = newAction - oldAction;
    const double logAcceptProb = logSampleRatio - currActionChange + prevActionChange;
    const double logu = log(rng_uniform(bd.seed, R.stream, move, TAG_ACCEPT, level));
    toAccept = logAcceptProb >= logu;
    prevActionChange = currActionChange;
    if (st) st->stages++;
    if (rec) {
      rec->old_action[stage] = oldAction;
      rec->new_action[stage] = newAction;
      rec->log_sample_ratio[stage] = logSampleRatio;
      rec->log_u[stage] = logu;
      rec->num_stages = stage + 1;
    }
  }
  if (toAccept) {
    // last Attempt (TablePermuteStage.cc:232-260): Rev->UpdateHTable on the NEW copy, then
    // CalcReverseProb (PermuteTableClass.cc:205-214) and the test
    double revT;
    if (len != 1) {
      std::vector<float> Hr = H;   // Rev's table is Forw's from InitBlock
      for (int i = 0; i < n; i++)
        for (int jC = 0; jC < len; jC++) {
          const int j = cyc.rep[jC];
          Hr[(size_t)i * n + j] = htable_entry(R, 1, first, i, j, slice1, slice2, logEps, fourLambdaBetaInv);
        }
      for (int iC = 0; iC < len; iC++) {
        const int i = cyc.rep[iC];
        for (int j = 0; j < n; j++)
          Hr[(size_t)i * n + j] = htable_entry(R, 1, first, i, j, slice1, slice2, logEps, fourLambdaBetaInv);
      }
      const double revNorm = boson_cycle_table(Hr.data(), n, Gamma, epsilon, nullptr);
      const double revNormInv = 1.0 / revNorm;
      revT = (Gamma[len - 1] * Gamma[len - 1]) / (cyc.P) * revNormInv;
    } else {
      revT = cyc.P * NormInv;
    }
    const double actionChange = -log(cyc.P / Gamma[len - 1]);
    const double pi_ratio = exp(-actionChange);
    const double Tratio = forwT / revT;
    const double acceptProb = std::min(1.0, pi_ratio / Tratio);
    const double psi = rng_uniform(bd.seed, R.stream, move, TAG_PERMUTE_ACCEPT, 0);
    toAccept = acceptProb > psi;
    if (rec) {
      rec->reverse_t = revT;
      rec->accept_prob = acceptProb;
      rec->psi = psi;
    }
  }
  if (toAccept)
    accept_copy(R, slice1, slice2, ptcls, len);
  else
    reject_copy(R, slice1, slice2, ptcls, len);
  if (rec) rec->accepted = toAccept ? 1 : 0;
  if (st) {
    st->moves++;
    if (toAccept) st->accepted++;
  }
}

int check_req(Ctx* c, const pimc_action_req& q, Replica** out) {
  Replica* R = get_rep(c, q.replica);
  if (!R) return fail(PIMC_ERR_ARG, "action request: bad replica id");
  if (q.mode != PIMC_OLDMODE && q.mode != PIMC_NEWMODE) return fail(PIMC_ERR_ARG, "action request: bad mode");
  if (q.level < 0 || q.level >= PIMC_MAX_LEVELS) return fail(PIMC_ERR_ARG, "action request: bad level");
  if (q.slice1 < 0 || q.slice2 > R->M || q.slice1 > q.slice2) return fail(PIMC_ERR_ARG, "action request: bad slice range");
  const int skip = 1 << q.level;
  if (q.slice2 > q.slice1 && q.slice1 + ((q.slice2 - q.slice1 + skip - 1) / skip) * skip > R->M)
    return fail(PIMC_ERR_ARG, "action request: slice range runs past the last slice at this level");
  if (q.nptcls < 1 || q.nptcls > PIMC_MAX_ACTIVE) return fail(PIMC_ERR_ARG, "action request: bad particle count");
  for (int i = 0; i < q.nptcls; i++)
    if (q.ptcls[i] < 0 || q.ptcls[i] >= R->N) return fail(PIMC_ERR_ARG, "action request: bad particle index");
  *out = R;
  return PIMC_OK;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
// A recorded sweep: the oracle replays it as ordinary bisection calls.
struct OracleSweep {
  std::vector<int32_t> replicas;
  pimc_bisection_desc desc{};
  int64_t first_move = 0, sweeps = 0;
  int32_t num_moves = 0, precision = 0;
  bool live = false;
};

struct pimc_ctx {
  Ctx c;
  double acc_ms = 0.0;   // time of the bisection calls since the last take
  pimc_bisection_plan plan{256, 1, 0, 1, 0.0};
  std::vector<OracleSweep> sweeps;
};

extern "C" {

int pimc_abi_version(void) { return PIMC_ABI_VERSION; }
const char* pimc_last_error(void) { return g_err.c_str(); }

int pimc_ctx_create(int32_t device, pimc_ctx** out) {
  (void)device;
  if (!out) return fail(PIMC_ERR_ARG, "ctx_create: out is NULL");
  *out = new pimc_ctx();
  return PIMC_OK;
}
int pimc_ctx_destroy(pimc_ctx* ctx) {
  delete ctx;
  return PIMC_OK;
}
int pimc_ctx_synchronize(pimc_ctx* ctx) { return ctx ? PIMC_OK : fail(PIMC_ERR_ARG, "NULL ctx"); }

int pimc_fit_create_david(pimc_ctx* ctx, const pimc_david_fit_desc* d, int32_t* fit_id) {
  if (!ctx || !d || !fit_id || !d->y) return fail(PIMC_ERR_ARG, "fit_create_david: NULL argument");
  if (d->num_points < 3 || d->num_tau < 1 || d->num_cols < 2) return fail(PIMC_ERR_ARG, "fit_create_david: bad sizes");
  int nt = d->num_terms;
  if (nt < 0 || (nt > 0 && d->num_cols < nt * (nt + 1) / 2 + nt + 2))
    return fail(PIMC_ERR_ARG, "fit_create_david: num_cols too small for num_terms");
  if (!(d->grid_start < d->grid_end)) return fail(PIMC_ERR_ARG, "fit_create_david: bad grid range");
  if (d->grid_type == PIMC_GRID_LOG && !(d->grid_start > 0)) return fail(PIMC_ERR_ARG, "fit_create_david: LOG grid needs start > 0");
  auto f = std::make_unique<DavidFit>();
  if (d->grid_type == PIMC_GRID_LINEAR)
    f->grid.init_linear(d->grid_start, d->grid_end, d->num_points);
  else if (d->grid_type == PIMC_GRID_LOG)
    f->grid.init_log_dm(d->grid_start, d->grid_end, d->num_points);
  else
    return fail(PIMC_ERR_ARG, "fit_create_david: bad grid type");
  f->num_tau = d->num_tau;
  f->ncols = d->num_cols;
  f->num_terms = nt;
  f->smallest_tau = d->smallest_tau;
  f->temp.resize(d->num_cols);
  std::vector<double> sd(d->num_cols, 5.0e30), ed(d->num_cols, 0.0);
  f->ukj.resize(d->num_tau);
  for (int l = 0; l < d->num_tau; l++)
    f->ukj[l].init(&f->grid, d->y + (size_t)l * d->num_points * d->num_cols, d->num_cols, sd.data(), ed.data());
  if (d->dy) {
    // dukj: column 0 the potential, the same end conditions (DavidPAClass.cc:840-877)
    f->dtemp.resize(d->num_cols);
    f->dukj.resize(d->num_tau);
    for (int l = 0; l < d->num_tau; l++)
      f->dukj[l].init(&f->grid, d->dy + (size_t)l * d->num_points * d->num_cols, d->num_cols, sd.data(), ed.data());
  }
  // TauPos (DavidPAClass.cc:859-868)
  f->taupos = -1;
  double tau = d->smallest_tau;
  for (int i = 0; i < d->num_tau; i++) {
    if (fabs(tau - d->desired_tau) < 1e-4) f->taupos = i;
    tau = tau * 2;
  }
  if (f->taupos < 0) return fail(PIMC_ERR_ARG, "fit_create_david: desired_tau matches no tau level");
  ctx->c.fits.push_back(std::move(f));
  *fit_id = (int32_t)ctx->c.fits.size() - 1;
  return PIMC_OK;
}

int pimc_fit_create_coulomb_bc(pimc_ctx* ctx, const pimc_coulomb_bc_fit_desc* d, int32_t* fit_id) {
  if (!ctx || !d || !fit_id || !d->umat) return fail(PIMC_ERR_ARG, "fit_create_coulomb_bc: NULL argument");
  if (d->q_points < 3 || d->t_points < 3 || d->num_betas < 1 || d->num_betas > PIMC_MAX_LEVELS)
    return fail(PIMC_ERR_ARG, "fit_create_coulomb_bc: bad sizes");
  auto f = std::make_unique<CoulombBCFit>();
  int rc = make_grid_checked(d->q_grid_type, d->q_start, d->q_end, d->q_points, f->qgrid);
  if (rc) return rc;
  rc = make_grid_checked(d->t_grid_type, d->t_start, d->t_end, d->t_points, f->tgrid);
  if (rc) return rc;
  f->SmallestBeta = d->smallest_beta;
  f->Z1Z2 = d->z1z2;
  f->Usplines.resize(d->num_betas);
  for (int b = 0; b < d->num_betas; b++)
    f->Usplines[b].init(&f->qgrid, &f->tgrid, d->umat + (size_t)b * d->q_points * d->t_points);
  if (d->dumat) {
    f->dUsplines.resize(d->num_betas);
    for (int b = 0; b < d->num_betas; b++)
      f->dUsplines[b].init(&f->qgrid, &f->tgrid, d->dumat + (size_t)b * d->q_points * d->t_points);
  }
  ctx->c.fits.push_back(std::move(f));
  *fit_id = (int32_t)ctx->c.fits.size() - 1;
  return PIMC_OK;
}

int pimc_replica_set_ulong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, int32_t num_levels,
                           double grid_start, double grid_end, int32_t num_points, const double* coefs,
                           const double* dcoefs) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || sa < 0 || sb < 0 || sa >= R->ns || sb >= R->ns || num_levels < 1 || num_points < 2 || !coefs ||
      !(grid_start < grid_end))
    return fail(PIMC_ERR_ARG, "replica_set_ulong: bad argument");
  QuinticTable t;
  t.set = true;
  t.grid.init_linear(grid_start, grid_end, num_points);
  t.levels = num_levels;
  t.c.assign(coefs, coefs + (size_t)num_levels * num_points * 6);
  R->ulong[sa * R->ns + sb] = t;
  R->ulong[sb * R->ns + sa] = t;
  if (dcoefs) {
    t.c.assign(dcoefs, dcoefs + (size_t)num_levels * num_points * 6);
    R->dulong[sa * R->ns + sb] = t;
    R->dulong[sb * R->ns + sa] = t;
  }
  return PIMC_OK;
}

int pimc_replica_set_vlong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, const double* vlong_k,
                           double vlong_r0, double vshort_k0, double vlong_k0, double grid_start, double grid_end,
                           int32_t num_points, const double* coefs) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || sa < 0 || sb < 0 || sa >= R->ns || sb >= R->ns || num_points < 2 || !coefs || !(grid_start < grid_end))
    return fail(PIMC_ERR_ARG, "replica_set_vlong: bad argument");
  if (R->nk > 0 && !vlong_k) return fail(PIMC_ERR_ARG, "replica_set_vlong: vlong_k is NULL");
  Replica::VlongTables t;
  t.set = true;
  t.spline.set = true;
  t.spline.grid.init_linear(grid_start, grid_end, num_points);
  t.spline.levels = 1;
  t.spline.c.assign(coefs, coefs + (size_t)num_points * 6);
  if (R->nk > 0) t.vk.assign(vlong_k, vlong_k + R->nk);
  t.r0 = vlong_r0;
  t.short_k0 = vshort_k0;
  t.long_k0 = vlong_k0;
  R->vlong[sa * R->ns + sb] = t;
  R->vlong[sb * R->ns + sa] = t;
  return PIMC_OK;
}

int pimc_fit_eval_u(pimc_ctx* ctx, int32_t fit_id, int32_t level, int64_t n, const double* q,
                    const double* z, const double* s2, int32_t precision, double* out) {
  (void)precision;
  if (!ctx || fit_id < 0 || fit_id >= (int)ctx->c.fits.size()) return fail(PIMC_ERR_ARG, "fit_eval_u: bad fit");
  const PairFit& f = *ctx->c.fits[fit_id];
  if (level < 0 || level >= f.num_levels()) return fail(PIMC_ERR_ARG, "fit_eval_u: level out of range");
  for (int64_t i = 0; i < n; i++) out[i] = f.U(q[i], z[i], s2[i], level);
  return PIMC_OK;
}

int pimc_replica_create(pimc_ctx* ctx, const pimc_replica_desc* d, int32_t* replica_id) {
  if (!ctx || !d || !replica_id) return fail(PIMC_ERR_ARG, "replica_create: NULL argument");
  if (d->num_particles < 1 || d->num_time_slices < 1 || d->num_species < 1 || d->num_species > PIMC_MAX_SPECIES)
    return fail(PIMC_ERR_ARG, "replica_create: bad sizes");
  auto R = std::make_unique<Replica>();
  R->N = d->num_particles;
  R->M = d->num_time_slices;
  R->nslices = R->M + 1;
  R->ns = d->num_species;
  R->num_images = d->num_images;
  R->tau = d->tau;
  for (int i = 0; i < 3; i++) {
    R->box[i] = d->box[i];
    R->boxinv[i] = 1.0 / d->box[i];
    R->isperiodic[i] = d->periodic[i] ? 1.0 : 0.0;
  }
  int tot = 0;
  for (int s = 0; s < R->ns; s++) {
    R->first.push_back(tot);
    R->count.push_back(d->species_count[s]);
    for (int i = 0; i < d->species_count[s]; i++) R->species_of.push_back(s);
    tot += d->species_count[s];
    R->lambda.push_back(d->lambda[s]);
  }
  if (tot != R->N) return fail(PIMC_ERR_ARG, "replica_create: species counts do not sum to num_particles");
  R->pair_fit.assign(d->pair_fit, d->pair_fit + R->ns * R->ns);
  for (int f : R->pair_fit)
    if (f < 0 || f >= (int)ctx->c.fits.size()) return fail(PIMC_ERR_ARG, "replica_create: bad pair fit id");
  size_t nb = (size_t)R->nslices * R->N * 3;
  R->path[0].assign(nb, 0.0);
  R->path[1].assign(nb, 0.0);
  R->perm[0].resize(R->N);
  R->perm[1].resize(R->N);
  for (int p = 0; p < R->N; p++) R->perm[0][p] = R->perm[1][p] = p;
  R->kcutoff = d->k_cutoff;
  R->use_long_range = d->use_long_range != 0;
  R->use_background = d->use_background != 0;
  R->stream = d->rng_stream;
  if (d->k_cutoff > 0) {
    R->long_range = true;
    R->setup_kvecs();
    R->rho[0].assign((size_t)R->nslices * R->ns * R->nk, cplx(0, 0));
    R->rho[1] = R->rho[0];
  }
  R->lr.resize(R->ns * R->ns);
  R->ulong.resize(R->ns * R->ns);
  R->dulong.resize(R->ns * R->ns);
  R->vlong.resize(R->ns * R->ns);
  ctx->c.reps.push_back(std::move(R));
  *replica_id = (int32_t)ctx->c.reps.size() - 1;
  return PIMC_OK;
}

int pimc_replica_num_kvecs(pimc_ctx* ctx, int32_t replica, int32_t* nk) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !nk) return fail(PIMC_ERR_ARG, "replica_num_kvecs: bad argument");
  *nk = R->nk;
  return PIMC_OK;
}
int pimc_replica_kvecs(pimc_ctx* ctx, int32_t replica, double* kvecs, int32_t* kidx) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R) return fail(PIMC_ERR_ARG, "replica_kvecs: bad replica");
  if (kvecs) memcpy(kvecs, R->kvecs.data(), R->kvecs.size() * sizeof(double));
  if (kidx) memcpy(kidx, R->kidx.data(), R->kidx.size() * sizeof(int32_t));
  return PIMC_OK;
}

int pimc_replica_set_long_range(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb,
                                int32_t num_levels, const double* ulong_k, const double* ulong_r0,
                                const double* ushort_k0, const double* ulong_k0, const double* dulong_k,
                                const double* dulong_r0, const double* dushort_k0, const double* dulong_k0) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || sa < 0 || sb < 0 || sa >= R->ns || sb >= R->ns || num_levels < 1 || !ulong_k || !ulong_r0 ||
      !ushort_k0 || !ulong_k0)
    return fail(PIMC_ERR_ARG, "replica_set_long_range: bad argument");
  const int nd = !!dulong_k + !!dulong_r0 + !!dushort_k0 + !!dulong_k0;
  if (nd != 0 && nd != 4) return fail(PIMC_ERR_ARG, "replica_set_long_range: give all four derivative tables or none");
  if (R->nk == 0) return fail(PIMC_ERR_STATE, "replica_set_long_range: replica has no k set");
  LongRangeTables t;
  t.set = true;
  t.num_levels = num_levels;
  t.ulong_k.assign(ulong_k, ulong_k + (size_t)num_levels * R->nk);
  t.ulong_r0.assign(ulong_r0, ulong_r0 + num_levels);
  t.ushort_k0.assign(ushort_k0, ushort_k0 + num_levels);
  t.ulong_k0.assign(ulong_k0, ulong_k0 + num_levels);
  if (nd == 4) {
    t.dset = true;
    t.dulong_k.assign(dulong_k, dulong_k + (size_t)num_levels * R->nk);
    t.dulong_r0.assign(dulong_r0, dulong_r0 + num_levels);
    t.dushort_k0.assign(dushort_k0, dushort_k0 + num_levels);
    t.dulong_k0.assign(dulong_k0, dulong_k0 + num_levels);
  }
  R->lr[sa * R->ns + sb] = t;
  R->lr[sb * R->ns + sa] = t;
  return PIMC_OK;
}

int pimc_path_set(pimc_ctx* ctx, int32_t replica, int32_t mode, const double* beads) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !beads || mode < 0 || mode > 2) return fail(PIMC_ERR_ARG, "path_set: bad argument");
  for (int m = 0; m < 2; m++)
    if (mode == 2 || mode == m) memcpy(R->path[m].data(), beads, R->path[m].size() * sizeof(double));
  return PIMC_OK;
}
int pimc_path_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* beads) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !beads || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "path_get: bad argument");
  memcpy(beads, R->path[mode].data(), R->path[mode].size() * sizeof(double));
  return PIMC_OK;
}
int pimc_path_set_beads(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t ptcl, int32_t slice1,
                        int32_t slice2, const double* xyz) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !xyz || mode < 0 || mode > 2 || ptcl < 0 || ptcl >= R->N || slice1 < 0 || slice2 > R->M || slice1 > slice2)
    return fail(PIMC_ERR_ARG, "path_set_beads: bad argument");
  for (int m = 0; m < 2; m++)
    if (mode == 2 || mode == m)
      for (int s = slice1; s <= slice2; s++) memcpy(R->P(m, s, ptcl), xyz + 3 * (s - slice1), 3 * sizeof(double));
  return PIMC_OK;
}
int pimc_permutation_get(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t* perm) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !perm || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "permutation_get: bad argument");
  memcpy(perm, R->perm[mode].data(), R->N * sizeof(int32_t));
  return PIMC_OK;
}

static int check_copy_args(Replica* R, int32_t slice1, int32_t slice2, const int32_t* ptcls, int32_t n) {
  if (!R || !ptcls || n < 0 || slice1 < 0 || slice2 > R->M || slice1 > slice2)
    return fail(PIMC_ERR_ARG, "accept/reject copy: bad argument");
  for (int i = 0; i < n; i++)
    if (ptcls[i] < 0 || ptcls[i] >= R->N) return fail(PIMC_ERR_ARG, "accept/reject copy: bad particle");
  return PIMC_OK;
}
int pimc_accept_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2, const int32_t* ptcls,
                     int32_t nptcls) {
  Replica* R = get_rep(&ctx->c, replica);
  int rc = check_copy_args(R, slice1, slice2, ptcls, nptcls);
  if (rc) return rc;
  accept_copy(*R, slice1, slice2, ptcls, nptcls);
  return PIMC_OK;
}
int pimc_reject_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2, const int32_t* ptcls,
                     int32_t nptcls) {
  Replica* R = get_rep(&ctx->c, replica);
  int rc = check_copy_args(R, slice1, slice2, ptcls, nptcls);
  if (rc) return rc;
  reject_copy(*R, slice1, slice2, ptcls, nptcls);
  return PIMC_OK;
}
int pimc_move_join(pimc_ctx* ctx, int32_t replica, int32_t old_join, int32_t new_join) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || old_join < 0 || new_join < 0 || old_join > R->M || new_join > R->M)
    return fail(PIMC_ERR_ARG, "move_join: bad argument");
  move_join(*R, old_join, new_join);
  R->join = new_join;
  return PIMC_OK;
}
int pimc_replica_join(pimc_ctx* ctx, int32_t replica, int32_t* join) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !join) return fail(PIMC_ERR_ARG, "replica_join: bad argument");
  *join = R->join;
  return PIMC_OK;
}
// PathClass::PutInBox (PathClass.h:897-921) / PutInBoxFast (PathClass.cc:1389-1407)
int pimc_put_in_box(pimc_ctx* ctx, int32_t replica, int32_t fast, int64_t n, double* v) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || n < 0 || (n > 0 && !v) || (fast != 0 && fast != 1)) return fail(PIMC_ERR_ARG, "put_in_box: bad argument");
  for (int64_t i = 0; i < n; i++) {
    if (fast)
      R->put_in_box_fast(v + 3 * i);
    else
      R->put_in_box(v + 3 * i);
  }
  return PIMC_OK;
}

// PermuteTableClass::ConstructHTable (PermuteTableClass.cc:122-150)
int pimc_permute_htable(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t species, int32_t slice1,
                        int32_t slice2, double epsilon, float* htable) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !htable || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "permute_htable: bad argument");
  if (species < 0 || species >= R->ns) return fail(PIMC_ERR_ARG, "permute_htable: bad species");
  if (slice1 < 0 || slice2 <= slice1 || slice2 >= R->nslices) return fail(PIMC_ERR_ARG, "permute_htable: bad slices");
  if (!(epsilon > 0.0)) return fail(PIMC_ERR_ARG, "permute_htable: epsilon must be positive");
  const double logEps = log(epsilon);
  const int firstPtcl = R->first[species];
  const double lambda = R->lambda[species];
  const double beta = R->tau * (double)(slice2 - slice1);
  const double fourLambdaBetaInv = 1.0 / (4.0 * lambda * beta);
  const int N = R->count[species];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      htable[(size_t)i * N + j] = htable_entry(*R, mode, firstPtcl, i, j, slice1, slice2, logEps, fourLambdaBetaInv);
  return PIMC_OK;
}

// TablePermuteStageClass::Read (TablePermuteStage.cc:142-166): PermuteType TABLE for `species`
int pimc_replica_set_permute(pimc_ctx* ctx, int32_t replica, int32_t species, const double* gamma,
                             double epsilon) {
  Replica* R = ctx ? get_rep(&ctx->c, replica) : nullptr;
  if (!R) return fail(PIMC_ERR_ARG, "replica_set_permute: bad replica");
  if (!gamma) {
    R->permute_on = false;
    return PIMC_OK;
  }
  if (species < 0 || species >= R->ns) return fail(PIMC_ERR_ARG, "replica_set_permute: bad species");
  if (!(epsilon > 0.0)) return fail(PIMC_ERR_ARG, "replica_set_permute: epsilon must be positive");
  for (int i = 0; i < 4; i++)
    if (!(gamma[i] > 0.0)) return fail(PIMC_ERR_ARG, "replica_set_permute: gamma must be positive");
  R->permute_on = true;
  R->permute_species = species;
  for (int i = 0; i < 4; i++) R->permute_gamma[i] = gamma[i];
  R->permute_epsilon = epsilon;
  return PIMC_OK;
}

// CycleClass::Apply (PermuteTableClass.cc:172-187), on the NEW copy
int pimc_cycle_apply(pimc_ctx* ctx, int32_t replica, int32_t first_ptcl, int32_t slice, int32_t length,
                     const int32_t* cycle) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !cycle || length < 1 || length > 4) return fail(PIMC_ERR_ARG, "cycle_apply: bad argument");
  if (slice < 0 || slice >= R->nslices || first_ptcl < 0) return fail(PIMC_ERR_ARG, "cycle_apply: bad slice or first_ptcl");
  for (int i = 0; i < length; i++) {
    if (cycle[i] < 0 || first_ptcl + cycle[i] >= R->N) return fail(PIMC_ERR_ARG, "cycle_apply: particle out of range");
    for (int k = 0; k < i; k++)
      if (cycle[k] == cycle[i]) return fail(PIMC_ERR_ARG, "cycle_apply: repeated particle");
  }
  double tempPos[3];
  memcpy(tempPos, R->P(1, slice, cycle[0] + first_ptcl), sizeof tempPos);
  const int tempPtcl = R->perm[1][cycle[0] + first_ptcl];
  for (int i = 0; i < length - 1; i++) {
    memcpy(R->P(1, slice, cycle[i] + first_ptcl), R->P(1, slice, cycle[i + 1] + first_ptcl), 3 * sizeof(double));
    R->perm[1][cycle[i] + first_ptcl] = R->perm[1][cycle[i + 1] + first_ptcl];
  }
  memcpy(R->P(1, slice, cycle[length - 1] + first_ptcl), tempPos, sizeof tempPos);
  R->perm[1][cycle[length - 1] + first_ptcl] = tempPtcl;
  return PIMC_OK;
}

int pimc_rho_k_update_all(pimc_ctx* ctx, int32_t replica) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R) return fail(PIMC_ERR_ARG, "rho_k_update_all: bad replica");
  if (R->nk == 0) return fail(PIMC_ERR_STATE, "rho_k_update_all: replica has no k set");
  R->update_rho_all();
  return PIMC_OK;
}
int pimc_rho_k_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* rho) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !rho || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "rho_k_get: bad argument");
  memcpy(rho, R->rho[mode].data(), R->rho[mode].size() * sizeof(cplx));
  return PIMC_OK;
}

int pimc_action_short_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                            double* out) {
  (void)precision;
  if (!ctx || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_short_range: NULL argument");
  for (int i = 0; i < nreq; i++) {
    Replica* R;
    int rc = check_req(&ctx->c, reqs[i], &R);
    if (rc) return rc;
    for (int f = 0; f < R->ns * R->ns; f++) {
      const PairFit& PA = *ctx->c.fits[R->pair_fit[f]];
      if (reqs[i].level >= PA.num_levels()) return fail(PIMC_ERR_ARG, "action_short_range: level beyond the fit's tau levels");
      if (PA.IsLongRange() && R->use_long_range && (!R->ulong[f].set || reqs[i].level >= R->ulong[f].levels))
        return fail(PIMC_ERR_STATE, "action_short_range: long-range pair fit without Ulong tables");
    }
    out[i] = short_range_action(&ctx->c, *R, reqs[i].mode, reqs[i].slice1, reqs[i].slice2, reqs[i].ptcls,
                                reqs[i].nptcls, reqs[i].level, nullptr);
  }
  return PIMC_OK;
}
// The no-allocation variant of the GPU library has nothing to allocate here: the same call.
int pimc_action_short_range_noalloc(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                                    double* out) {
  if (nreq > PIMC_NOALLOC_MAX_REQS)
    return fail(PIMC_ERR_ARG, "action_short_range_noalloc: more than PIMC_NOALLOC_MAX_REQS requests");
  return pimc_action_short_range(ctx, reqs, nreq, precision, out);
}
int pimc_action_kinetic(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, double* out) {
  if (!ctx || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_kinetic: NULL argument");
  for (int i = 0; i < nreq; i++) {
    Replica* R;
    int rc = check_req(&ctx->c, reqs[i], &R);
    if (rc) return rc;
    out[i] = kinetic_action(*R, reqs[i].mode, reqs[i].slice1, reqs[i].slice2, reqs[i].ptcls, reqs[i].nptcls,
                            reqs[i].level);
  }
  return PIMC_OK;
}
int pimc_action_long_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                           double* out) {
  (void)precision;
  if (!ctx || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_long_range: NULL argument");
  for (int i = 0; i < nreq; i++) {
    Replica* R;
    int rc = check_req(&ctx->c, reqs[i], &R);
    if (rc) return rc;
    if (R->nk == 0) return fail(PIMC_ERR_STATE, "action_long_range: replica has no k set");
    out[i] = long_range_action(*R, reqs[i].mode, reqs[i].slice1, reqs[i].slice2, reqs[i].ptcls, reqs[i].nptcls,
                               reqs[i].level, nullptr);
  }
  return PIMC_OK;
}

int pimc_dist_disp(pimc_ctx* ctx, int32_t replica, int32_t mode, int64_t n, const int32_t* idx, double* out,
                   int32_t* images) {
  Replica* R = get_rep(&ctx->c, replica);
  if (!R || !idx || !out || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "dist_disp: bad argument");
  for (int64_t i = 0; i < n; i++) {
    const int32_t* q = idx + 4 * i;
    if (q[0] < 0 || q[0] > R->M || q[1] < 0 || q[1] > R->M || q[2] < 0 || q[2] >= R->N || q[3] < 0 || q[3] >= R->N)
      return fail(PIMC_ERR_ARG, "dist_disp: index out of range");
    double* o = out + 8 * i;
    double nimg[3], mimg[3];
    R->dist_disp(mode, q[0], q[1], q[2], q[3], o[0], o[1], o + 2, o + 5, nimg, mimg);
    if (images)
      for (int d = 0; d < 3; d++) {
        images[6 * i + d] = (int32_t)nimg[d];
        images[6 * i + 3 + d] = (int32_t)mimg[d];
      }
  }
  return PIMC_OK;
}

// KineticClass::d_dBeta (KineticClass.cc:187-256)
static double kinetic_dbeta(Replica& R, int mode, int slice1, int slice2, int level) {
  double spring = 0.0;
  double levelTau = ldexp(R.tau, level);
  int skip = 1 << level;
  for (int ptcl = 0; ptcl < R.N; ptcl++) {
    double lambda = R.lambda[R.species_of[ptcl]];
    if (lambda != 0.0) {
      double FourLambdaTauInv = 1.0 / (4.0 * lambda * levelTau);
      for (int slice = slice1; slice < slice2; slice += skip) {
        spring += (0.5 * 3) / levelTau;
        double vel[3];
        R.velocity(mode, slice, slice + skip, ptcl, vel);
        double Z = 1.0;
        double GaussSum[3] = {0.0, 0.0, 0.0}, numSum[3] = {0.0, 0.0, 0.0};
        for (int dim = 0; dim < 3; dim++) {
          for (int image = -R.num_images; image <= R.num_images; image++) {
            double dist = vel[dim] + (double)image * R.box[dim];
            double d2overFLT = dist * dist * FourLambdaTauInv;
            double expPart = exp(-d2overFLT);
            GaussSum[dim] += expPart;
            numSum[dim] += -d2overFLT / levelTau * expPart;
          }
          Z *= GaussSum[dim];
        }
        double scalarnumSum = 0.0;
        for (int dim = 0; dim < 3; dim++) {
          double numProd = 1.0;
          for (int dim2 = 0; dim2 < 3; dim2++) numProd *= (dim2 != dim) ? GaussSum[dim2] : numSum[dim2];
          scalarnumSum += numProd;
        }
        spring += scalarnumSum / Z;
      }
    }
  }
  return spring;
}

// ShortRangeClass::d_dBeta (ShortRangeClass.cc:256-295)
static double short_range_dbeta(Ctx* c, Replica& R, int mode, int slice1, int slice2, int level) {
  int skip = 1 << level;
  double dU = 0.0;
  for (int ptcl1 = 0; ptcl1 < R.N; ptcl1++) {
    int species1 = R.species_of[ptcl1];
    for (int ptcl2 = 0; ptcl2 < ptcl1; ptcl2++) {
      int pr = species1 * R.ns + R.species_of[ptcl2];
      const PairFit& pa = *c->fits[R.pair_fit[pr]];
      for (int slice = slice1; slice < slice2; slice += skip) {
        double r[3], rp[3], rmag, rpmag;
        R.dist_disp(mode, slice, slice + skip, ptcl1, ptcl2, rmag, rpmag, r, rp, nullptr, nullptr);
        double d0 = r[0] - rp[0], d1 = r[1] - rp[1], d2 = r[2] - rp[2];
        double s2 = d0 * d0 + d1 * d1 + d2 * d2;
        double q = 0.5 * (rmag + rpmag);
        double z = (rmag - rpmag);
        dU += pa.dU(q, z, s2, level);
        if (pa.IsLongRange() && R.use_long_range)
          dU -= 0.5 * (R.dulong[pr].eval(level, rmag) + R.dulong[pr].eval(level, rpmag));
      }
    }
  }
  return dU;
}

// LongRangeClass::d_dBeta (LongRangeClass.cc:374-442) with rho_k recomputed from the
// positions (CalcRho_ks_Fast) into a scratch copy.
static double long_range_dbeta(Replica& R, int mode, int slice1, int slice2, int level) {
  double homo = 0.0, hetero = 0.0, background = 0.0, k0Homo = 0.0, k0Hetero = 0.0;
  int skip = (1 << level);
  std::vector<cplx> saved = R.rho[mode];
  for (int slice = slice1; slice <= slice2; slice += skip) {
    double factor = ((slice == slice1) || (slice == slice2)) ? 0.5 : 1.0;
    for (int species = 0; species < R.ns; species++) {
      R.calc_rho_fast(mode, slice, species);
      const LongRangeTables& PA = R.lr[species * R.ns + species];
      if (PA.set) {
        for (int ki = 0; ki < R.nk; ki++) {
          double rhok2 = std::norm(R.R(mode, slice, species, ki));
          homo += factor * 0.5 * 2.0 * rhok2 * PA.dulong_k[(size_t)level * R.nk + ki];
        }
        int N = R.count[species];
        homo -= factor * 0.5 * N * PA.dulong_r0[level];
        background -= factor * 0.5 * N * N * PA.dushort_k0[level];
        k0Homo += factor * 0.5 * N * N * PA.dulong_k0[level];
      }
    }
    for (int species1 = 0; species1 < R.ns; species1++)
      for (int species2 = species1 + 1; species2 < R.ns; species2++) {
        const LongRangeTables& PA = R.lr[species1 * R.ns + species2];
        if (PA.set) {
          for (int ki = 0; ki < R.nk; ki++) {
            const cplx a = R.R(mode, slice, species1, ki), b = R.R(mode, slice, species2, ki);
            double rhorho = a.real() * b.real() + a.imag() * b.imag();
            hetero += factor * 2.0 * rhorho * PA.dulong_k[(size_t)level * R.nk + ki];
          }
          int N1 = R.count[species1], N2 = R.count[species2];
          background -= factor * N1 * N2 * PA.dushort_k0[level];
          k0Hetero += factor * N1 * N2 * PA.dulong_k0[level];
        }
      }
  }
  R.rho[mode] = saved;
  double dU = homo + hetero;
  if (R.use_background)
    dU += background;
  else
    dU += (k0Homo + k0Hetero);
  return dU;
}

// ShortRangePotClass::V(slice) (ShortRangePotClass.cc:27-47) with DistDisp(slice, p1, p2)
// (PathClass.h:638-660): V of the pair's fit (column 0 of a DavidPA table, Z1Z2/r of a
// Coulomb fit), minus Vlong(r) for a long-ranged fit when long range is used.
static double short_range_v(Ctx* c, Replica& R, int mode, int slice) {
  double val = 0.0;
  for (int ptcl1 = 0; ptcl1 < R.N; ptcl1++) {
    int species1 = R.species_of[ptcl1];
    for (int ptcl2 = 0; ptcl2 < ptcl1; ptcl2++) {
      const double* a = R.P(mode, slice, ptcl1);
      const double* b = R.P(mode, slice, ptcl2);
      double disp[3];
      for (int d = 0; d < 3; d++) {
        disp[d] = b[d] - a[d];
        double n = nearbyint(disp[d] * R.boxinv[d]);
        disp[d] -= n * R.isperiodic[d] * R.box[d];
      }
      double dist = sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
      const int f = species1 * R.ns + R.species_of[ptcl2];
      const PairFit& pa = *c->fits[R.pair_fit[f]];
      val += pa.V(dist);
      if (pa.IsLongRange() && R.use_long_range) val -= R.vlong[f].spline.eval(0, dist);
    }
  }
  return val;
}

// LongRangePotClass::V (LongRangePotClass.cc:32-85) with the stored rho_k of the slice.
static double long_range_v(Replica& R, int mode, int slice) {
  double homo = 0.0, hetero = 0.0, background = 0.0, k0Homo = 0.0, k0Hetero = 0.0;
  for (int species = 0; species < R.ns; species++) {
    const Replica::VlongTables& T = R.vlong[species * R.ns + species];
    if (!T.set) continue;   // no long-range potential for this pair (tables zero in the reference)
    for (int ki = 0; ki < R.nk; ki++) {
      const cplx z = R.R(mode, slice, species, ki);
      homo += 0.5 * 2.0 * (z.real() * z.real() + z.imag() * z.imag()) * T.vk[ki];
    }
    const int N = R.count[species];
    homo -= 0.5 * N * T.r0;
    background -= 0.5 * N * N * T.short_k0;
    k0Homo += 0.5 * N * N * T.long_k0;
  }
  for (int species1 = 0; species1 < R.ns; species1++)
    for (int species2 = species1 + 1; species2 < R.ns; species2++) {
      const Replica::VlongTables& T = R.vlong[species1 * R.ns + species2];
      if (!T.set) continue;
      for (int ki = 0; ki < R.nk; ki++) {
        const cplx a = R.R(mode, slice, species1, ki), b = R.R(mode, slice, species2, ki);
        hetero += 2.0 * (a.real() * b.real() + a.imag() * b.imag()) * T.vk[ki];
      }
      const int N1 = R.count[species1], N2 = R.count[species2];
      background -= N1 * N2 * T.short_k0;
      k0Hetero += N1 * N2 * T.long_k0;
    }
  double Vlong = homo + hetero;
  if (R.use_background)
    Vlong += background;
  else
    Vlong += (k0Homo + k0Hetero);
  return Vlong;
}

// ActionBaseClass::d_dBeta(slice1, slice2, level) of one action (ActionBase.h:57).
int pimc_action_d_dbeta(pimc_ctx* ctx, int32_t action, const pimc_range_req* reqs, int32_t nreq, double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad argument");
  if (action != PIMC_ACTION_SHORT_RANGE && action != PIMC_ACTION_KINETIC && action != PIMC_ACTION_LONG_RANGE)
    return fail(PIMC_ERR_ARG, "action_d_dbeta: bad action");
  for (int i = 0; i < nreq; i++) {
    const pimc_range_req& q = reqs[i];
    Replica* R = get_rep(&ctx->c, q.replica);
    if (!R) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad replica id");
    if (q.mode != PIMC_OLDMODE && q.mode != PIMC_NEWMODE) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad mode");
    if (q.level < 0 || q.level >= PIMC_MAX_LEVELS) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad level");
    const int skip = 1 << q.level;
    if (q.slice1 < 0 || q.slice1 > q.slice2 || q.slice1 + ((q.slice2 - q.slice1 + skip - 1) / skip) * skip > R->M)
      return fail(PIMC_ERR_ARG, "action_d_dbeta: bad slice range");
    if (action == PIMC_ACTION_LONG_RANGE && R->nk == 0) return fail(PIMC_ERR_STATE, "action_d_dbeta: replica has no k set");
    if (action != PIMC_ACTION_KINETIC)
      for (int f = 0; f < R->ns * R->ns; f++) {
        const PairFit& pa = *ctx->c.fits[R->pair_fit[f]];
        if (!pa.has_dU()) return fail(PIMC_ERR_STATE, "action_d_dbeta: a pair fit has no beta-derivative table");
        if (pa.IsLongRange() && R->use_long_range && (!R->dulong[f].set || R->dulong[f].levels <= q.level))
          return fail(PIMC_ERR_STATE, "action_d_dbeta: long-range pair fit without dUlong tables");
        if (R->long_range && R->use_long_range && R->lr[f].set && !R->lr[f].dset)
          return fail(PIMC_ERR_STATE, "action_d_dbeta: long-range tables without beta derivatives");
      }
  }
  for (int i = 0; i < nreq; i++) {
    const pimc_range_req& q = reqs[i];
    Replica& R = *get_rep(&ctx->c, q.replica);
    if (action == PIMC_ACTION_KINETIC)
      out[i] = kinetic_dbeta(R, q.mode, q.slice1, q.slice2, q.level);
    else if (action == PIMC_ACTION_SHORT_RANGE)
      out[i] = short_range_dbeta(&ctx->c, R, q.mode, q.slice1, q.slice2, q.level);
    else
      out[i] = (R.long_range && R.use_long_range) ? long_range_dbeta(R, q.mode, q.slice1, q.slice2, q.level) : 0.0;
  }
  return PIMC_OK;
}

int pimc_estimator_energy(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, int32_t mode, pimc_energy* out) {
  if (!ctx || nrep < 0 || (nrep > 0 && (!replicas || !out)) || mode < 0 || mode > 1)
    return fail(PIMC_ERR_ARG, "estimator_energy: bad argument");
  for (int r = 0; r < nrep; r++) {
    Replica* R = get_rep(&ctx->c, replicas[r]);
    if (!R) return fail(PIMC_ERR_ARG, "estimator_energy: bad replica");
    for (int f = 0; f < R->ns * R->ns; f++) {
      const PairFit& pa = *ctx->c.fits[R->pair_fit[f]];
      if (!pa.has_dU()) return fail(PIMC_ERR_STATE, "estimator_energy: a pair fit has no beta-derivative table");
      if (pa.IsLongRange() && R->use_long_range && (!R->dulong[f].set || R->dulong[f].levels < 1))
        return fail(PIMC_ERR_STATE, "estimator_energy: long-range pair fit without dUlong tables");
      if (R->long_range && R->use_long_range && R->lr[f].set && !R->lr[f].dset)
        return fail(PIMC_ERR_STATE, "estimator_energy: long-range tables without beta derivatives");
    }
  }
  for (int r = 0; r < nrep; r++) {
    Replica& R = *get_rep(&ctx->c, replicas[r]);
    const int M = R.M;
    pimc_energy e{};
    e.kinetic = kinetic_dbeta(R, mode, 0, M, 0);
    e.du_short = short_range_dbeta(&ctx->c, R, mode, 0, M, 0);
    e.du_long = (R.long_range && R.use_long_range) ? long_range_dbeta(R, mode, 0, M, 0) : 0.0;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    bool v_known = true, vlong_known = true;
    for (int f = 0; f < R.ns * R.ns; f++) {
      const PairFit& pa = *ctx->c.fits[R.pair_fit[f]];
      if (pa.IsLongRange() && R.use_long_range && !R.vlong[f].set) v_known = false;
      if (R.long_range && R.use_long_range && R.lr[f].set && !R.vlong[f].set) vlong_known = false;
    }
    double v = 0.0, vl = 0.0;
    for (int slice = 0; slice <= M; slice++) {
      double factor = ((slice == 0) || (slice == M)) ? 0.5 : 1.0;
      if (v_known) v += factor * short_range_v(&ctx->c, R, mode, slice);
      if (R.long_range && R.use_long_range && vlong_known) vl += factor * long_range_v(R, mode, slice);
    }
    e.v_short = v_known ? v : nan;
    e.v_long = vlong_known ? vl : nan;
    e.total = e.kinetic + e.du_short + e.du_long;
    out[r] = e;
  }
  return PIMC_OK;
}

// PairCorrelationClass::Accumulate (ObservableCorrelation.cc:156-221)
int pimc_estimator_pair_correlation(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t s1, int32_t s2,
                                    double start, double end, int32_t num_points, int64_t* hist) {
  Replica* Rp = ctx ? get_rep(&ctx->c, replica) : nullptr;
  if (!Rp || !hist || mode < 0 || mode > 1 || num_points < 2 || !(start < end) || s1 < 0 || s2 < 0 ||
      s1 >= Rp->ns || s2 >= Rp->ns)
    return fail(PIMC_ERR_ARG, "estimator_pair_correlation: bad argument");
  Replica& R = *Rp;
  Grid grid;
  grid.init_linear(start, end, num_points);
  auto add = [&](int slice, int p1, int p2) {
    const double* a = R.P(mode, slice, p1);
    const double* b = R.P(mode, slice, p2);
    double disp[3];
    for (int d = 0; d < 3; d++) {
      disp[d] = b[d] - a[d];
      double n = nearbyint(disp[d] * R.boxinv[d]);
      disp[d] -= n * R.isperiodic[d] * R.box[d];
    }
    double dist = sqrt(disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
    if (dist < grid.end && dist >= grid.start) {
      int index = grid.reverse_map(dist);
      if (index >= 0 && index < num_points - 1) hist[index]++;
    }
  };
  const int f1 = R.first[s1], l1 = R.first[s1] + R.count[s1] - 1;
  const int f2 = R.first[s2], l2 = R.first[s2] + R.count[s2] - 1;
  for (int slice = 0; slice < R.nslices - 1; slice++) {
    if (s1 == s2) {
      for (int p1 = f1; p1 <= l1; p1++)
        for (int p2 = p1 + 1; p2 <= l1; p2++) add(slice, p1, p2);
    } else {
      for (int p1 = f1; p1 <= l1; p1++)
        for (int p2 = f2; p2 <= l2; p2++) add(slice, p1, p2);
    }
  }
  return PIMC_OK;
}

// The GPU library's launch rules for TABLE permutation moves, mirrored so both backends refuse
// the same calls: every listed replica moves `species` with TABLE or none does, and a TABLE
// replica has no long-range action.
static int check_table_launch(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, int species, int r,
                              const char* who) {
  const Replica* R = get_rep(&ctx->c, replicas[r]);
  const bool table = R->permute_on && R->permute_species == species;
  const Replica* R0 = get_rep(&ctx->c, replicas[0]);
  if (R0 && table != (R0->permute_on && R0->permute_species == species))
    return fail(PIMC_ERR_UNSUPPORTED, std::string(who) + ": replicas with and without TABLE permutations in one call");
  if (table && has_long_range(*R))
    return fail(PIMC_ERR_UNSUPPORTED, std::string(who) + ": TABLE permutation moves with the long-range action");
  (void)nrep;
  return PIMC_OK;
}

int pimc_bisection_run(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                       int64_t first_move, int32_t num_moves, int32_t precision, pimc_move_record* records,
                       pimc_bisection_stats* stats) {
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "bisection_run: bad precision");
  if (!ctx || !replicas || !desc || nrep < 0 || num_moves < 0 || first_move < 0 ||
      first_move + num_moves > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "bisection_run: bad argument");
  if (desc->num_levels < 1 || desc->num_levels > PIMC_MAX_LEVELS || desc->lowest_level < 0 ||
      desc->lowest_level >= desc->num_levels)
    return fail(PIMC_ERR_ARG, "bisection_run: bad levels");
  for (int r = 0; r < nrep; r++) {
    Replica* R = get_rep(&ctx->c, replicas[r]);
    if (!R) return fail(PIMC_ERR_ARG, "bisection_run: bad replica");
    if ((1 << desc->num_levels) >= R->nslices) return fail(PIMC_ERR_ARG, "bisection_run: too few slices for num_levels");
    if (desc->species < 0 || desc->species >= R->ns) return fail(PIMC_ERR_ARG, "bisection_run: bad species");
    int rc = check_table_launch(ctx, replicas, nrep, desc->species, r, "bisection_run");
    if (rc) return rc;
    for (int s = 0; s < R->ns; s++)
      for (int t = 0; t < R->ns; t++) {
        const PairFit& f = *ctx->c.fits[R->pair_fit[s * R->ns + t]];
        if (desc->num_levels > f.num_levels())
          return fail(PIMC_ERR_ARG, "bisection_run: pair fit has too few tau levels");
        const QuinticTable& ul = R->ulong[s * R->ns + t];
        if (f.IsLongRange() && R->use_long_range && (!ul.set || desc->num_levels > ul.levels))
          return fail(PIMC_ERR_STATE, "bisection_run: long-range pair fit without Ulong tables");
      }
  }
  pimc_bisection_stats st;
  memset(&st, 0, sizeof st);
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < nrep; r++) {
    Replica& R = *get_rep(&ctx->c, replicas[r]);
    pimc_bisection_stats one;
    memset(&one, 0, sizeof one);
    for (int m = 0; m < num_moves; m++) {
      pimc_move_record* rec = records ? records + (size_t)r * num_moves + m : nullptr;
      bisection_move(&ctx->c, R, *desc, (uint32_t)(first_move + m), rec, &one);
    }
    st.moves += one.moves;
    st.accepted += one.accepted;
    st.stages += one.stages;
    st.pair_evals += one.pair_evals;
    st.k_evals += one.k_evals;
    R.acc.moves += one.moves;
    R.acc.accepted += one.accepted;
    R.acc.stages += one.stages;
    R.acc.pair_evals += one.pair_evals;
    R.acc.k_evals += one.k_evals;
  }
  st.device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ctx->acc_ms += st.device_ms;
  if (stats) *stats = st;
  return PIMC_OK;
}

// Launch shapes mean nothing on the CPU: the oracle validates and stores them.
int pimc_bisection_set_plan(pimc_ctx* ctx, const pimc_bisection_plan* plan) {
  if (!ctx || !plan) return fail(PIMC_ERR_ARG, "bisection_set_plan: NULL argument");
  if (plan->block_threads != 256 && plan->block_threads != 512)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: block_threads must be 256 or 512");
  if (plan->link_split != 1 && plan->link_split != 2 && plan->link_split != 4)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: link_split must be 1, 2 or 4");
  if (plan->cluster < 0 || plan->cluster > 32) return fail(PIMC_ERR_ARG, "bisection_set_plan: cluster must be 0..32");
  if (plan->cluster > 1 && plan->persistent)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: a cluster plan is not persistent");
  ctx->plan = *plan;
  ctx->plan.persistent = plan->persistent ? 1 : 0;
  ctx->plan.cluster = plan->cluster > 1 ? plan->cluster : 1;
  return PIMC_OK;
}

int pimc_bisection_get_plan(pimc_ctx* ctx, pimc_bisection_plan* plan) {
  if (!ctx || !plan) return fail(PIMC_ERR_ARG, "bisection_get_plan: NULL argument");
  *plan = ctx->plan;
  return PIMC_OK;
}

int pimc_bisection_tune(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                        int64_t first_move, int32_t trial_moves, int32_t precision, pimc_bisection_plan* out) {
  if (!ctx || !desc || nrep <= 0 || !replicas || trial_moves <= 0 || first_move < 0 ||
      first_move + trial_moves > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "bisection_tune: bad argument");
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "bisection_tune: bad precision");
  for (int r = 0; r < nrep; r++) {
    Replica* R = get_rep(&ctx->c, replicas[r]);
    if (!R) return fail(PIMC_ERR_ARG, "bisection_tune: bad replica");
  }
  // the checks of a run (a TABLE list must be all TABLE, and so on), without moving anything
  const pimc_bisection_plan caller = ctx->plan;
  ctx->plan = pimc_bisection_plan{256, 1, 0, 1, 0.0};
  const int rc = pimc_bisection_run(ctx, replicas, nrep, desc, first_move, 0, precision, nullptr, nullptr);
  if (rc) {
    ctx->plan = caller;
    return rc;
  }
  if (out) *out = ctx->plan;
  return PIMC_OK;
}

int pimc_sweep_graph_create(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                            int64_t first_move, int32_t num_moves, int32_t precision, int32_t* graph_id) {
  if (!ctx || !desc || nrep <= 0 || !replicas || num_moves <= 0 || first_move < 0 || !graph_id ||
      first_move > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "sweep_graph_create: bad argument");
  // the same checks as a run, without moving anything
  int rc = pimc_bisection_run(ctx, replicas, nrep, desc, first_move, 0, precision, nullptr, nullptr);
  if (rc) return rc;
  for (int r = 0; r < nrep; r++) {
    const Replica* R = get_rep(&ctx->c, replicas[r]);
    if (R->permute_on && R->permute_species == desc->species)
      return fail(PIMC_ERR_UNSUPPORTED, "sweep_graph_create: TABLE permutation moves are not captured in graphs");
  }
  OracleSweep g;
  g.replicas.assign(replicas, replicas + nrep);
  g.desc = *desc;
  g.first_move = first_move;
  g.num_moves = num_moves;
  g.precision = precision;
  g.live = true;
  ctx->sweeps.push_back(g);
  *graph_id = (int32_t)ctx->sweeps.size() - 1;
  return PIMC_OK;
}

int pimc_sweep_graph_launch(pimc_ctx* ctx, int32_t graph_id, int32_t num_sweeps) {
  if (!ctx || graph_id < 0 || graph_id >= (int)ctx->sweeps.size() || !ctx->sweeps[graph_id].live || num_sweeps < 0)
    return fail(PIMC_ERR_ARG, "sweep_graph_launch: bad argument");
  OracleSweep& g = ctx->sweeps[graph_id];
  if (g.first_move + (int64_t)g.num_moves * (g.sweeps + num_sweeps) > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "sweep_graph_launch: move numbers beyond 2^32");
  for (int i = 0; i < num_sweeps; i++) {
    int rc = pimc_bisection_run(ctx, g.replicas.data(), (int32_t)g.replicas.size(), &g.desc,
                                g.first_move + g.sweeps * g.num_moves, g.num_moves, g.precision, nullptr, nullptr);
    if (rc) return rc;
    g.sweeps++;
  }
  return PIMC_OK;
}

int pimc_sweep_graph_destroy(pimc_ctx* ctx, int32_t graph_id) {
  if (!ctx || graph_id < 0 || graph_id >= (int)ctx->sweeps.size() || !ctx->sweeps[graph_id].live)
    return fail(PIMC_ERR_ARG, "sweep_graph_destroy: bad argument");
  ctx->sweeps[graph_id].live = false;
  return PIMC_OK;
}

int pimc_bisection_stats_take(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, pimc_bisection_stats* stats) {
  if (!ctx || !stats || nrep < 0 || (nrep > 0 && !replicas)) return fail(PIMC_ERR_ARG, "bisection_stats_take: bad argument");
  for (int r = 0; r < nrep; r++)
    if (!get_rep(&ctx->c, replicas[r])) return fail(PIMC_ERR_ARG, "bisection_stats_take: bad replica");
  memset(stats, 0, sizeof *stats);
  stats->device_ms = ctx->acc_ms;
  ctx->acc_ms = 0.0;
  for (int r = 0; r < nrep; r++) {
    Replica& R = *get_rep(&ctx->c, replicas[r]);
    stats->moves += R.acc.moves;
    stats->accepted += R.acc.accepted;
    stats->stages += R.acc.stages;
    stats->pair_evals += R.acc.pair_evals;
    stats->k_evals += R.acc.k_evals;
    R.acc = pimc_bisection_stats{};
  }
  return PIMC_OK;
}

// ---------------- oracle-only hooks ----------------
void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox(ctr, key, out); }
double oracle_uniform(uint32_t w0, uint32_t w1) { return uniform53(w0, w1); }
int oracle_gaussian_pair(uint64_t seed, uint32_t stream, uint32_t move, int32_t level, uint32_t pair, double* g) {
  return gaussian_pair(seed, stream, move, level, pair, g);
}

static int make_grid(int32_t type, double start, double end, int32_t npoints, Grid& g) {
  if (npoints < 2 || !(start < end)) return fail(PIMC_ERR_ARG, "grid: bad arguments");
  if (type == PIMC_GRID_LINEAR)
    g.init_linear(start, end, npoints);
  else if (type == PIMC_GRID_LOG && start > 0)
    g.init_log_dm(start, end, npoints);
  else
    return fail(PIMC_ERR_ARG, "grid: bad type");
  return PIMC_OK;
}
int oracle_reverse_map(int32_t type, double start, double end, int32_t npoints, int32_t literal, int64_t n,
                       const double* r, int32_t* out) {
  Grid g;
  int rc = make_grid(type, start, end, npoints, g);
  if (rc) return rc;
  for (int64_t i = 0; i < n; i++) out[i] = literal ? g.reverse_map_literal(r[i]) : g.reverse_map(r[i]);
  return PIMC_OK;
}
int oracle_grid_points(int32_t type, double start, double end, int32_t npoints, double* x) {
  Grid g;
  int rc = make_grid(type, start, end, npoints, g);
  if (rc) return rc;
  memcpy(x, g.x.data(), npoints * sizeof(double));
  return PIMC_OK;
}
int oracle_multicubic(int32_t type, double start, double end, int32_t npoints, int32_t ncols, const double* y,
                      const double* start_deriv, const double* end_deriv, double* d2y, int32_t col, int64_t n,
                      const double* x, double* out) {
  Grid g;
  int rc = make_grid(type, start, end, npoints, g);
  if (rc) return rc;
  MultiCubicSpline s;
  s.init(&g, y, ncols, start_deriv, end_deriv);
  if (d2y) memcpy(d2y, s.d2y.data(), s.d2y.size() * sizeof(double));
  for (int64_t i = 0; i < n; i++) {
    if (col >= 0)
      out[i] = s.eval(col, x[i]);
    else
      s.eval_all(x[i], out + i * ncols);
  }
  return PIMC_OK;
}
int oracle_cycle_table(const float* htable, int32_t n, const double* gamma, double epsilon, int32_t cap,
                       int32_t* len, int32_t* rep, float* P, float* C, float* norm, double xi, int32_t* found) {
  std::vector<CycleEntry> table;
  *norm = boson_cycle_table(htable, n, gamma, epsilon, &table);
  for (int e = 0; e < (int)table.size() && e < cap; e++) {
    len[e] = table[e].len;
    for (int q = 0; q < 4; q++) rep[4 * e + q] = q < table[e].len ? table[e].rep[q] : -1;
    P[e] = table[e].P;
    C[e] = table[e].C;
  }
  if (xi >= 0.0 && found) *found = find_entry(table, (double)*norm, xi);
  return (int)table.size();
}

int oracle_bisection_sample(void* ctx, int32_t replica, int32_t ptcl, int32_t slice1, int32_t num_levels,
                            int32_t level, uint64_t seed, uint32_t move, double* log_sample_ratio) {
  pimc_ctx* c = (pimc_ctx*)ctx;
  Replica* R = get_rep(&c->c, replica);
  if (!R) return fail(PIMC_ERR_ARG, "bisection_sample: bad replica");
  int p[1] = {ptcl};
  *log_sample_ratio = bisection_sample(*R, slice1, num_levels, level, p, 1, seed, move);
  return PIMC_OK;
}

}  // extern "C"
