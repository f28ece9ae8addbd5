// pimc_estimators.hip -- the thermodynamic energy and pair-correlation estimators.
//
//   kinetic   KineticClass::d_dBeta          PIMC++/src/Actions/KineticClass.cc:187-256
//   du_short  ShortRangeClass::d_dBeta       PIMC++/src/Actions/ShortRangeClass.cc:256-295
//             DavidPAClass::dU / calcUsqz    Common/PairAction/DavidPAClass.cc:115-126, 291-463
//             PAcoulombBCFitClass::dU        Common/PairAction/PAcoulombBCFit.cc:178-192
//   v_short   ShortRangePotClass::V          PIMC++/src/Actions/ShortRangePotClass.cc:27-47
//   du_long   LongRangeClass::d_dBeta        PIMC++/src/Actions/LongRangeClass.cc:374-442
//   g(r)      PairCorrelationClass::Accumulate  PIMC++/src/Observables/ObservableCorrelation.cc:156-221
//
// They run over the whole path and are off the timed path: FP64 throughout, tables
// read from global memory.  Every block writes its partial sums; the host adds
// them in block order, so results do not depend on scheduling.
#include "pimc_kernels.hpp"
#include "pimc_launch.hpp"

namespace pimc {

// Interval and local coordinate of x in a fit's r grid (clamped to [0, nint-1]).
__device__ inline int fit_low(const FitDev& F, double x, double& b) {
  const int l = spline_low(F, F.x, x);
  const d2v g = F.geo[l];
  b = (x - g.x) * g.y;
  return l;
}

// calcUsqz's dU (DavidPAClass.cc:291-463) for bisection level `level`: r, r'
// clamped into the grid; dU = (du00(r)+du00(r'))/2 + (V(r)+V(r'))/2 + sum of the
// dukj off-diagonal terms where s > 0 and q < End.
template <int NT>
__device__ inline double david_du(const FitDev& F, int level, double q, double z, double s2) {
  const d4v* E = F.e64 + (size_t)(level + F.taupos) * F.nint * F.ecol;
  const double rmin = F.start;
  double r = q + 0.5 * z, rp = q - 0.5 * z;
  if (r > F.end) r = F.end;
  if (rp > F.end) rp = F.end;
  if (rp < rmin) rp = rmin;
  if (r < rmin) r = rmin;
  double br, bp;
  const int lr = fit_low(F, r, br), lp = fit_low(F, rp, bp);
  const double V = 0.5 * (horner64(E[lr * F.ecol], br) + horner64(E[lp * F.ecol], bp));
  double dU = 0.5 * (horner64(E[lr * F.ecol + 1], br) + horner64(E[lp * F.ecol + 1], bp));
  dU += V;
  if constexpr (NT > 0) {
    const double s = sqrt(s2);
    if (s > 0.0 && q < F.end) {
      const double zsquared = z * z, ssquared = s * s, ssquaredinverse = 1.0 / ssquared;
      double bq;
      const int lq = fit_low(F, q, bq);
      double Sto2k = ssquared;
      int col = 2;
#pragma unroll
      for (int k = 1; k <= NT; k++) {
        double Zto2j = 1.0, currS = Sto2k;
#pragma unroll
        for (int j = 0; j <= k; j++, col++) {
          dU += horner64(E[lq * F.ecol + col], bq) * Zto2j * currS;
          Zto2j *= zsquared;
          currS = currS * ssquaredinverse;
        }
        Sto2k = Sto2k * ssquared;
      }
    }
  }
  return dU;
}

__device__ inline double bc_du(const FitDev& F, int level, double q, double s2) {
  const GridDev Q = qgrid_of(F);
  const d4v* tab = F.dbc + (size_t)level * F.ng * F.nt;
  if (q <= (F.end * 1.0000001)) {
    if (q == 0) return bicubic(F, Q, tab, 0.0, 0.0);
    return bicubic(F, Q, tab, q, sqrt(s2) / (2.0 * q));
  }
  return F.z1z2 / q;
}

// DavidPAClass::V: column 0 of the first stored level at r clamped to the grid.
__device__ inline double david_v(const FitDev& F, double r) {
  if (r > F.end) r = F.end;
  if (r < F.start) r = F.start;
  double b;
  const int l = fit_low(F, r, b);
  return horner64(F.e64[l * F.ecol], b);
}

// Pair p of the N(N-1)/2 pairs (ptcl1 > ptcl2), in the reference's loop order.
__device__ inline void pair_of(long long p, int& p1, int& p2) {
  long long i = (long long)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
  while (i * (i - 1) / 2 > p) i--;
  while ((i + 1) * i / 2 <= p) i++;
  p1 = (int)i;
  p2 = (int)(p - i * (i - 1) / 2);
}

// Nearest-image distance of one slice (DistDisp(slice, p1, p2), PathClass.h:638-660).
__device__ inline double slice_dist(const RepCore& R, const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double disp = b[d] - a[d];
    const double n = rint(disp * R.boxinv[d]);
    disp -= n * R.isper[d] * R.box[d];
    d2 += disp * disp;
  }
  return sqrt(d2);
}

// dU of every pair over the links [slice1, slice2) at `level`, and V over the slices
// slice1, slice1 + 2^level, .., slice2 (ends weighted 1/2), of copy `mode`.
// partials[block] = {du, v}.
template <int NT>
__global__ void __launch_bounds__(kBlock) k_energy_pairs(const ReplicaDev* reps, const FitDev* fits, int rep, int mode,
                                                         int slice1, int slice2, int level, double* partials) {
  __shared__ double scratch_raw[kWaves];
  PIMC_LDS double* scratch = (PIMC_LDS double*)scratch_raw;
  const ReplicaDev* RD = reps + rep;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int skip = 1 << level;
  const long long npairs = (long long)R.N * (R.N - 1) / 2;
  double du = 0.0, v = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npairs;
       p += (long long)gridDim.x * blockDim.x) {
    int p1, p2;
    pair_of(p, p1, p2);
    int s1 = 0, s2 = 0;
    for (int s = 1; s < R.ns; s++) {
      if (p1 >= S.first[s]) s1 = s;
      if (p2 >= S.first[s]) s2 = s;
    }
    const int pr = s1 * R.ns + s2;
    const FitDev& F = fits[S.fit[pr]];
    const double* dul = (F.kind == 1 && R.use_long_range && R.dulong)
                            ? R.dulong + ((size_t)pr * R.ulong_levels + level) * R.ulong_n * 6
                            : nullptr;
    // ShortRangePotClass::V: V(r) (column 0 of a DavidPA fit, Z1Z2/r of a Coulomb fit), minus
    // Vlong(r) for a long-ranged fit when long range is used
    const double* vl = (F.kind == 1 && R.use_long_range && ((R.vlong_mask >> pr) & 1u))
                           ? R.vlong + (size_t)pr * R.ulong_n * 6
                           : nullptr;
    auto vpair = [&](double r) {
      double val = F.kind == 0 ? david_v(F, r) : F.z1z2 / r;
      if (vl) val -= ulong_eval(R, vl, r);
      return val;
    };
    double a1[3], a2[3];
    load_bead(R, mode, slice1, p1, a1);
    load_bead(R, mode, slice1, p2, a2);
    v += 0.5 * vpair(slice_dist(R, a1, a2));
    for (int slice = slice1; slice < slice2; slice += skip) {
      double b1[3], b2[3];
      load_bead(R, mode, slice + skip, p1, b1);
      load_bead(R, mode, slice + skip, p2, b2);
      double da[3], db[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double n, m;
        dist_disp_1(a2[d] - a1[d], b2[d] - b1[d], R.box[d], R.boxinv[d], R.isper[d], da[d], db[d], n, m);
      }
      const double rmag = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
      const double rpmag = sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]);
      const double e0 = da[0] - db[0], e1 = da[1] - db[1], e2 = da[2] - db[2];
      const double s2v = e0 * e0 + e1 * e1 + e2 * e2;
      const double q = 0.5 * (rmag + rpmag), z = rmag - rpmag;
      du += F.kind == 1 ? bc_du(F, level, q, s2v) : david_du<NT>(F, level, q, z, s2v);
      if (dul) du -= 0.5 * (ulong_eval(R, dul, rmag) + ulong_eval(R, dul, rpmag));
      v += (slice + skip >= slice2 ? 0.5 : 1.0) * vpair(slice_dist(R, b1, b2));
#pragma unroll
      for (int d = 0; d < 3; d++) {
        a1[d] = b1[d];
        a2[d] = b2[d];
      }
    }
  }
  const double tdu = block_sum(du, scratch);
  const double tv = block_sum(v, scratch);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = tdu;
    partials[2 * blockIdx.x + 1] = tv;
  }
}

// KineticClass::d_dBeta over every particle and the links [slice1, slice2) at `level` of copy `mode`.
__global__ void __launch_bounds__(kBlock) k_energy_kinetic(const ReplicaDev* reps, int rep, int mode, int slice1,
                                                           int slice2, int level, double* partials) {
  __shared__ double scratch_raw[kWaves];
  PIMC_LDS double* scratch = (PIMC_LDS double*)scratch_raw;
  const ReplicaDev* RD = reps + rep;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int skip = 1 << level;
  const int nlinks = (slice2 - slice1 + skip - 1) / skip;
  const double levelTau = ldexp(R.tau, level);
  const long long nitems = (long long)R.N * nlinks;
  double spring = 0.0;
  for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < nitems;
       it += (long long)gridDim.x * blockDim.x) {
    const int ptcl = (int)(it / nlinks), slice = slice1 + (int)(it % nlinks) * skip;
    int sp = 0;
    for (int s = 1; s < R.ns; s++)
      if (ptcl >= S.first[s]) sp = s;
    const double lambda = S.lambda[sp];
    if (lambda == 0.0) continue;
    const double FourLambdaTauInv = 1.0 / (4.0 * lambda * levelTau);
    spring += (0.5 * 3) / levelTau;
    double a[3], b[3];
    load_bead(R, mode, slice, ptcl, a);
    load_bead(R, mode, slice + skip, ptcl, b);
    double GaussSum[3], numSum[3], Z = 1.0;
#pragma unroll
    for (int dim = 0; dim < 3; dim++) {
      const double vel = put_in_box_1(b[dim] - a[dim], R.box[dim], R.boxinv[dim], R.isper[dim]);
      GaussSum[dim] = 0.0;
      numSum[dim] = 0.0;
      for (int image = -R.num_images; image <= R.num_images; image++) {
        const double dist = vel + (double)image * R.box[dim];
        const double d2overFLT = dist * dist * FourLambdaTauInv;
        const double expPart = exp(-d2overFLT);
        GaussSum[dim] += expPart;
        numSum[dim] += -d2overFLT / levelTau * expPart;
      }
      Z *= GaussSum[dim];
    }
    double scalarnumSum = 0.0;
#pragma unroll
    for (int dim = 0; dim < 3; dim++) {
      double numProd = 1.0;
#pragma unroll
      for (int dim2 = 0; dim2 < 3; dim2++) numProd *= (dim2 != dim) ? GaussSum[dim2] : numSum[dim2];
      scalarnumSum += numProd;
    }
    spring += scalarnumSum / Z;
  }
  const double t = block_sum(spring, scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// LongRangeClass::d_dBeta: one block per slice slice1 + b 2^level (b = blockIdx.x) up to
// slice2, rho_k of that slice from rho ([b][species][k]).  partials[b] = {homo + hetero,
// background, k0} with the slice's factor, then the same three sums of the long-range
// potential.
__global__ void __launch_bounds__(kBlock) k_energy_long_range(const ReplicaDev* reps, int rep, int slice1, int slice2,
                                                              int level, const d2v* __restrict__ rho_in,
                                                              double* partials) {
  __shared__ double scratch_raw[kWaves];
  PIMC_LDS double* scratch = (PIMC_LDS double*)scratch_raw;
  const ReplicaDev* RD = reps + rep;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int b = blockIdx.x;
  const int slice = slice1 + b * (1 << level);
  const double factor = (slice == slice1 || slice == slice2) ? 0.5 : 1.0;
  const d2v* rho = rho_in + (size_t)b * R.ns * R.nk;
  const int L = R.lr_levels;
  double part = 0.0, vpart = 0.0;   // d_dBeta and LongRangePotClass::V (LongRangePotClass.cc:32-85)
  for (int k = threadIdx.x; k < R.nk; k += blockDim.x) {
    for (int s = 0; s < R.ns; s++) {
      const int pr = s * R.ns + s;
      const d2v a = rho[(size_t)s * R.nk + k];
      if ((R.lr_mask >> pr) & 1u)
        part += factor * 0.5 * 2.0 * (a.x * a.x + a.y * a.y) * R.dulong_k[((size_t)pr * L + level) * R.nk + k];
      if ((R.vlong_mask >> pr) & 1u) vpart += factor * 0.5 * 2.0 * (a.x * a.x + a.y * a.y) * R.vlong_k[(size_t)pr * R.nk + k];
    }
    for (int s1 = 0; s1 < R.ns; s1++)
      for (int s2 = s1 + 1; s2 < R.ns; s2++) {
        const int pr = s1 * R.ns + s2;
        const d2v a = rho[(size_t)s1 * R.nk + k], c = rho[(size_t)s2 * R.nk + k];
        if ((R.lr_mask >> pr) & 1u)
          part += factor * 2.0 * (a.x * c.x + a.y * c.y) * R.dulong_k[((size_t)pr * L + level) * R.nk + k];
        if ((R.vlong_mask >> pr) & 1u) vpart += factor * 2.0 * (a.x * c.x + a.y * c.y) * R.vlong_k[(size_t)pr * R.nk + k];
      }
  }
  const double t = block_sum(part, scratch);
  const double tv = block_sum(vpart, scratch);
  if (threadIdx.x == 0) {
    double homo = 0.0, background = 0.0, k0 = 0.0;
    for (int s = 0; s < R.ns; s++) {
      const int pr = s * R.ns + s;
      if (!((R.lr_mask >> pr) & 1u)) continue;
      const int N = S.count[s];
      homo -= factor * 0.5 * N * R.dlr_const[(size_t)(pr * 3 + 0) * L + level];
      background -= factor * 0.5 * N * N * R.dlr_const[(size_t)(pr * 3 + 1) * L + level];
      k0 += factor * 0.5 * N * N * R.dlr_const[(size_t)(pr * 3 + 2) * L + level];
    }
    for (int s1 = 0; s1 < R.ns; s1++)
      for (int s2 = s1 + 1; s2 < R.ns; s2++) {
        const int pr = s1 * R.ns + s2;
        if (!((R.lr_mask >> pr) & 1u)) continue;
        const int N1 = S.count[s1], N2 = S.count[s2];
        background -= factor * N1 * N2 * R.dlr_const[(size_t)(pr * 3 + 1) * L + level];
        k0 += factor * N1 * N2 * R.dlr_const[(size_t)(pr * 3 + 2) * L + level];
      }
    partials[6 * b] = t + homo;
    partials[6 * b + 1] = background;
    partials[6 * b + 2] = k0;
    // the potential's Madelung, background and k = 0 terms
    double vhomo = 0.0, vbackground = 0.0, vk0 = 0.0;
    for (int s = 0; s < R.ns; s++) {
      const int pr = s * R.ns + s;
      if (!((R.vlong_mask >> pr) & 1u)) continue;
      const int N = S.count[s];
      vhomo -= factor * 0.5 * N * R.vlr_const[(size_t)pr * 3 + 0];
      vbackground -= factor * 0.5 * N * N * R.vlr_const[(size_t)pr * 3 + 1];
      vk0 += factor * 0.5 * N * N * R.vlr_const[(size_t)pr * 3 + 2];
    }
    for (int s1 = 0; s1 < R.ns; s1++)
      for (int s2 = s1 + 1; s2 < R.ns; s2++) {
        const int pr = s1 * R.ns + s2;
        if (!((R.vlong_mask >> pr) & 1u)) continue;
        const int N1 = S.count[s1], N2 = S.count[s2];
        vbackground -= factor * N1 * N2 * R.vlr_const[(size_t)pr * 3 + 1];
        vk0 += factor * N1 * N2 * R.vlr_const[(size_t)pr * 3 + 2];
      }
    partials[6 * b + 3] = tv + vhomo;
    partials[6 * b + 4] = vbackground;
    partials[6 * b + 5] = vk0;
  }
}

// PairCorrelationClass::Accumulate into hist (64-bit counters).  Each block counts its items
// in a histogram of 32-bit counters in LDS (lds_hist: nbins <= kLdsBins, and fewer than 2^32
// items per block) and adds it to hist once, so the global counters see one atomic per bin
// and block instead of one per pair; otherwise every count goes to hist directly.  The items
// (slice, pair) are walked with an incremental index, without a 64-bit division per item.
constexpr int kLdsBins = 8192;
__global__ void __launch_bounds__(kBlock) k_pair_correlation(const ReplicaDev* reps, int rep, int mode, int f1,
                                                             int n1, int f2, int n2, int same, double start,
                                                             double deltainv, double end, int nbins, int lds_hist,
                                                             unsigned long long* hist) {
  extern __shared__ unsigned int lh[];   // [nbins] when lds_hist
  const ReplicaDev* RD = reps + rep;
  const RepCore R = load_uniform(&RD->core);
  const int M = R.nslices - 1;
  const long long npairs = same ? (long long)n1 * (n1 - 1) / 2 : (long long)n1 * n2;
  if (lds_hist) {
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) lh[b] = 0u;
    __syncthreads();
  }
  if (npairs > 0) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long sq = stride / npairs, sr = stride % npairs;
    const long long it0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long slice = it0 / npairs, p = it0 % npairs;
    for (; slice < M; slice += sq, p += sr) {
      if (p >= npairs) {
        p -= npairs;
        slice++;
        if (slice >= M) break;
      }
      int p1, p2;
      if (same) {
        int hi, lo;
        pair_of(p, hi, lo);   // hi > lo
        p1 = f1 + lo;
        p2 = f1 + hi;
      } else {
        p1 = f1 + (int)(p / n2);
        p2 = f2 + (int)(p % n2);
      }
      double a[3], b[3];
      load_bead(R, mode, (int)slice, p1, a);
      load_bead(R, mode, (int)slice, p2, b);
      const double dist = slice_dist(R, a, b);
      if (dist < end && dist >= start) {
        int index;
        {
#pragma clang fp contract(off)
          index = clamp_to_int(rint((dist - start) * deltainv - 0.5), nbins + 1);
        }
        if (index >= 0 && index < nbins) {
          if (lds_hist)
            atomicAdd(lh + index, 1u);
          else
            atomicAdd(hist + index, 1ull);
        }
      }
    }
  }
  if (lds_hist) {
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
      if (lh[b]) atomicAdd(hist + b, (unsigned long long)lh[b]);
  }
}

hipError_t launch_energy(int num_terms, const ReplicaDev* reps, const FitDev* fits, int rep, int mode, int slice1,
                         int slice2, int level, int pair_blocks, int kin_blocks, int lr_slices, const d2v* lr_rho,
                         double* pair_partials, double* kin_partials, double* lr_partials, hipStream_t s) {
  if (pair_blocks > 0) {
#define PIMC_PAIRS(NT_) k_energy_pairs<NT_><<<pair_blocks, kBlock, 0, s>>>(reps, fits, rep, mode, slice1, slice2, level, pair_partials)
    switch (num_terms) {
      case 0: PIMC_PAIRS(0); break;
      case 1: PIMC_PAIRS(1); break;
      case 2: PIMC_PAIRS(2); break;
      default: PIMC_PAIRS(3); break;
    }
#undef PIMC_PAIRS
  }
  if (kin_blocks > 0) k_energy_kinetic<<<kin_blocks, kBlock, 0, s>>>(reps, rep, mode, slice1, slice2, level, kin_partials);
  if (lr_slices > 0)
    k_energy_long_range<<<lr_slices, kBlock, 0, s>>>(reps, rep, slice1, slice2, level, lr_rho, lr_partials);
  return hipGetLastError();
}

hipError_t launch_pair_correlation(const ReplicaDev* reps, int rep, int mode, int f1, int n1, int f2, int n2,
                                   int same, double start, double deltainv, double end, int nbins,
                                   unsigned long long* hist, long long nitems, int blocks, hipStream_t s) {
  // a block counts at most ceil(nitems / threads) items per thread
  const long long threads = (long long)blocks * kBlock;
  const int lds_hist = nbins <= kLdsBins && (nitems + threads - 1) / threads * kBlock < (1LL << 32);
  k_pair_correlation<<<blocks, kBlock, lds_hist ? (size_t)nbins * sizeof(unsigned int) : 0, s>>>(
      reps, rep, mode, f1, n1, f2, n2, same, start, deltainv, end, nbins, lds_hist, hist);
  return hipGetLastError();
}

}  // namespace pimc
