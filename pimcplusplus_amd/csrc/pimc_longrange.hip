// pimc_longrange.hip -- rho_k kernels and LongRangeClass::SingleAction.
//
//   k_rho_all   PathClass::UpdateRho_ks()        PathClass.cc:709-723 (+ CalcRho_ks_Fast :616-647);
//               CalcRho_ks_Fast of chosen slices of one copy into a scratch array (d_dBeta)
//   k_lr_req    LongRangeClass::SingleAction     LongRangeClass.cc:273-371, with
//               PathClass::UpdateRho_ks(s1, s2, changed, level) PathClass.cc:649-707 in NEWMODE
// rho_k is [slice][species][k] complex double; a thread owns one k and adds
// the particles in the reference's order, so every rho_k sum runs in the same
// order as the CPU path.
#include <algorithm>
#include <unordered_map>

#include "pimc_kernels.hpp"
#include "pimc_launch.hpp"

namespace pimc {

constexpr int kRhoChunk = 16;   // particles whose phase tables are staged together

// rho_k of slice slice1 + blockIdx.x * skip and species blockIdx.y from copy `mode`: into
// out[blockIdx.x][species][k] when out is given, else into both mirrored copies.
__global__ void __launch_bounds__(kBlock) k_rho_all(const ReplicaDev* __restrict__ reps, int rep, int mode, int slice1,
                                                    int skip, d2v* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  PIMC_LDS d2v* C = (PIMC_LDS d2v*)smem_raw;
  const ReplicaDev* RD = reps + rep;
  const RepCore R = load_uniform(&RD->core);
  const int slice = slice1 + (int)blockIdx.x * skip, s = blockIdx.y;
  const double* path = R.path(mode);
  const int k = blockIdx.z * blockDim.x + threadIdx.x;
  const int first = RD->sp.first[s], end = first + RD->sp.count[s];
  const int stride = kstride(R);
  const int o1 = koff_of(R, 1), o2 = koff_of(R, 2);
  int ki[3] = {0, 0, 0};
  if (k < R.nk) {
    ki[0] = R.kidx[4 * k];
    ki[1] = R.kidx[4 * k + 1];
    ki[2] = R.kidx[4 * k + 2];
  }
  double re = 0.0, im = 0.0;
  for (int p0 = first; p0 < end; p0 += kRhoChunk) {
    const int np = min(kRhoChunk, end - p0);
    __syncthreads();
    for (int t = threadIdx.x; t < np * 3; t += blockDim.x) {
      const int p = p0 + t / 3, d = t % 3;
      const double x = path[((size_t)slice * 3 + d) * R.N + p];
      const double kb = R.kbox[0] * (double)(d == 0) + R.kbox[1] * (double)(d == 1) + R.kbox[2] * (double)(d == 2);
      phase_row<double>(x * kb, kmax_of(R, d), C + (size_t)(t / 3) * stride + koff_of(R, d));
    }
    __syncthreads();
    if (k < R.nk)
      for (int pp = 0; pp < np; pp++) {
        double tr, ti;
        kterm<double>(C + (size_t)pp * stride, o1, o2, ki, tr, ti);
        re += tr;
        im += ti;
      }
  }
  if (k < R.nk) {
    if (out) {
      out[((size_t)blockIdx.x * R.ns + s) * R.nk + k] = d2v{re, im};
    } else {
      const size_t at = ((size_t)slice * R.ns + s) * R.nk + k;
      R.rho0[at] = d2v{re, im};
      R.rho1[at] = d2v{re, im};
    }
  }
}

void launch_rho_all(const ReplicaDev* reps, int rep, int nslices, int ns, int nk, int stride, hipStream_t s) {
  dim3 grid(nslices, ns, (nk + kBlock - 1) / kBlock);
  k_rho_all<<<grid, kBlock, (size_t)kRhoChunk * stride * 16, s>>>(reps, rep, 0, 0, 1, nullptr);
}

void launch_rho_slices(const ReplicaDev* reps, int rep, int mode, int slice1, int skip, int nsl, int ns, int nk,
                       int stride, d2v* out, hipStream_t s) {
  dim3 grid(nsl, ns, (nk + kBlock - 1) / kBlock);
  k_rho_all<<<grid, kBlock, (size_t)kRhoChunk * stride * 16, s>>>(reps, rep, mode, slice1, skip, out);
}

// LongRangeClass::SingleAction for request req0 + blockIdx.y over the k-vectors
// [blockIdx.x kLrChunk, (blockIdx.x + 1) kLrChunk) of its replica, one k per thread.
// NEWMODE writes every (slice, species, k) of rho_k NEW over the range once: OLD plus, for
// the moved particles in order, the new phase minus the old one (PathClass::UpdateRho_ks
// (slice1, slice2, changed, level), PathClass.cc:649-707), and the action terms are formed
// from the value written.  So rho_k is streamed once: 16 B read per (k, slice, species) in
// OLDMODE, 16 B read and 16 B written in NEWMODE.  Each thread walks the slices with the
// loads of the next three (NEWMODE launches) or four slices in flight.  The phase rows of
// the moved beads (NEW and OLD copies) are staged in LDS for as many slices as kLrPhaseLds
// holds.
// partials[request][blockIdx.x] holds the block's sum; block 0 adds the k-independent terms.
// The requests of one launch belong to different replicas.
constexpr int kLrChunk = 256;
constexpr size_t kLrPhaseLds = 24 * 1024;
constexpr size_t kLrScratch = 64;   // block-sum scratch ahead of the phase rows

// The k-dependent terms of one slice (LongRangeClass.cc:296-330): homologous
// factor |rho_s|^2 u_ss and heterologous factor 2 Re(rho_a rho_b*) u_ab.
template <int NSP>
__device__ inline double lr_slice_terms(const RepCore& R, const d2v* v, const double* u, double factor) {
  double part = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; s++)
    if (s < R.ns && ((R.lr_mask >> (s * R.ns + s)) & 1u))
      part += factor * 0.5 * 2.0 * (v[s].x * v[s].x + v[s].y * v[s].y) * u[s * NSP + s];
#pragma unroll
  for (int s1 = 0; s1 < NSP; s1++)
#pragma unroll
    for (int s2 = s1 + 1; s2 < NSP; s2++)
      if (s2 < R.ns && ((R.lr_mask >> (s1 * R.ns + s2)) & 1u))
        part += factor * 2.0 * (v[s1].x * v[s2].x + v[s1].y * v[s2].y) * u[s1 * NSP + s2];
  return part;
}

// One phase row of the request kernel's staging, out of line: its sincospi and recursion
// would otherwise set the register count of the whole streaming kernel.
template <typename PT>
__device__ __noinline__ void lr_stage_row(const RepCore& R, const double* row, int n, int d, PIMC_LDS PhaseRow<PT>* C) {
  const double x[3] = {row[0], row[n], row[2 * n]};
  phase_rows_lane<PT>(R, x, d, C);
}

// UPD: the launch holds NEWMODE requests (an OLDMODE-only launch needs no phase rows, holds
// fewer registers and keeps more loads in flight).
template <bool MIXED, int NSP, bool UPD>
__global__ void __launch_bounds__(kBlock) k_lr_req(const ReplicaDev* __restrict__ reps, const ReqDev* __restrict__ reqs,
                                                   int req0, int kblocks, double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  typedef typename std::conditional<MIXED, float, double>::type PT;
  PIMC_LDS double* scratch = (PIMC_LDS double*)smem_raw;
  PIMC_LDS PhaseRow<PT>* C = (PIMC_LDS PhaseRow<PT>*)(smem_raw + kLrScratch);
  const int r = req0 + (int)blockIdx.y;
  const ReqDev Q = reqs[r];
  const PIMC_GLOBAL int* qp = (const PIMC_GLOBAL int*)reqs[r].ptcls;   // indexed at run time: not from Q
  const ReplicaDev* RD = reps + Q.replica;
  const RepCore& R = RD->core;   // read in place (scalar loads): a register copy of R spills
  const RepSpecies& S = RD->sp;
  const int stride = kstride(R), o1 = koff_of(R, 1), o2 = koff_of(R, 2);
  const int skip = 1 << Q.level;
  const int nsl = (Q.slice2 - Q.slice1) / skip + 1;
  const bool upd = UPD && Q.mode == 1;
  const int np = upd ? Q.nptcls : 0;
  const int k = (int)blockIdx.x * kLrChunk + (int)threadIdx.x;
  const bool active = k < R.nk;
  const int rows = (int)(kLrPhaseLds / ((size_t)stride * sizeof(PhaseRow<PT>)));
  const int cs = upd ? max(1, min(nsl, rows / max(1, 2 * np))) : nsl;   // slices per staged chunk
  const size_t sstride = (size_t)R.ns * R.nk;                             // distance between slices
  // global-address-space views: global rather than flat loads and stores
  const PIMC_GLOBAL d2v* rho_old = (const PIMC_GLOBAL d2v*)R.rho0;
  PIMC_GLOBAL d2v* rho_new = (PIMC_GLOBAL d2v*)R.rho1;
  int ki[3] = {0, 0, 0};
  double u[NSP * NSP];
#pragma unroll
  for (int i = 0; i < NSP * NSP; i++) u[i] = 0.0;
  int psp = 0;   // species of moved particle pi in bits 2pi, 2pi+1
  double part = 0.0;
  for (int c0 = 0; c0 < nsl; c0 += cs) {
    const int nc = min(cs, nsl - c0);
    // rho_k OLD of slice j in v[0] and, in flight, of slices j + 1 .. j + PF; the first PF
    // are issued before the phase rows are staged
    constexpr int PF = UPD ? 3 : 4;
    d2v v[PF + 1][NSP];
    auto load = [&](int j, d2v* v) {
#pragma unroll
      for (int s = 0; s < NSP; s++)
        v[s] = (active && j < nc && s < R.ns)
                   ? rho_old[(size_t)(Q.slice1 + (c0 + j) * skip) * sstride + (size_t)s * R.nk + k]
                   : d2v{0.0, 0.0};
    };
#pragma unroll
    for (int b = 0; b < PF; b++) load(b, v[b]);
    if (c0 == 0) {   // this thread's k index and u_k, the moved particles' species
      if (active) {
        const PIMC_GLOBAL int* kx = (const PIMC_GLOBAL int*)R.kidx + 4 * k;
        ki[0] = kx[0];
        ki[1] = kx[1];
        ki[2] = kx[2];
        const PIMC_GLOBAL double* uk = (const PIMC_GLOBAL double*)R.ulong_k;
#pragma unroll
        for (int s1 = 0; s1 < NSP; s1++)
#pragma unroll
          for (int s2 = s1; s2 < NSP; s2++)
            if (s2 < R.ns && ((R.lr_mask >> (s1 * R.ns + s2)) & 1u))
              u[s1 * NSP + s2] = uk[((size_t)(s1 * R.ns + s2) * R.lr_levels + Q.level) * R.nk + k];
      }
      for (int pi = 0; pi < np; pi++) {
        const int p = qp[pi];
        int sp = 0;
        while (sp + 1 < R.ns && p >= S.first[sp + 1]) sp++;
        psp |= sp << (2 * pi);
      }
    }
    if (UPD && upd) {
      __syncthreads();
      for (int i = threadIdx.x; i < nc * np * 6; i += blockDim.x) {
        const int j = i / (np * 6), pi = (i / 6) % np, c = (i % 6) / 3, d = i % 3;
        // copy NEW (c = 0) or OLD (c = 1), chosen arithmetically: a select between fields of
        // the register-resident R becomes a load from a private copy of R
        const uintptr_t pa = (uintptr_t)R.path1 * (uintptr_t)(c == 0) + (uintptr_t)R.path0 * (uintptr_t)(c != 0);
        const double* row = (const double*)pa + (size_t)(Q.slice1 + (c0 + j) * skip) * 3 * R.N + qp[pi];
        lr_stage_row<PT>(R, row, R.N, d, C + (size_t)((j * np + pi) * 2 + c) * stride);
      }
      __syncthreads();
    }
    if (!active) continue;
    for (int j = 0; j < nc; j++) {
      load(j + PF, v[PF]);
      d2v* v0 = v[0];
      const int slice = Q.slice1 + (c0 + j) * skip;
      if (UPD && upd) {
        for (int pi = 0; pi < np; pi++) {
          const int sp = (psp >> (2 * pi)) & 3;
          PT nr, ni, orr, oi;
          kterm<PT>(C + (size_t)((j * np + pi) * 2) * stride, o1, o2, ki, nr, ni);
          kterm<PT>(C + (size_t)((j * np + pi) * 2 + 1) * stride, o1, o2, ki, orr, oi);
#pragma unroll
          for (int s = 0; s < NSP; s++)
            if (s == sp) {
              v0[s].x += (double)nr;
              v0[s].y += (double)ni;
              v0[s].x -= (double)orr;
              v0[s].y -= (double)oi;
            }
        }
#pragma unroll
        for (int s = 0; s < NSP; s++)
          if (s < R.ns) {   // streamed out: not read again by this launch
            PIMC_GLOBAL double* at = (PIMC_GLOBAL double*)(rho_new + (size_t)slice * sstride + (size_t)s * R.nk + k);
            __builtin_nontemporal_store(v0[s].x, at);
            __builtin_nontemporal_store(v0[s].y, at + 1);
          }
      }
      const double factor = (slice == Q.slice1 || slice == Q.slice2) ? 0.5 : 1.0;
      part += lr_slice_terms<NSP>(R, v0, u, factor);
#pragma unroll
      for (int b = 0; b < PF; b++)
#pragma unroll
        for (int s = 0; s < NSP; s++) v[b][s] = v[b + 1][s];
    }
  }
  const double ksum = block_sum(part, scratch);
  if (threadIdx.x == 0) {
    double U = ksum;
    if (blockIdx.x == 0) {
      // the k-independent terms are linear in the slice factor: added once with their sum
      double fsum = 0.0, homo = 0.0, background = 0.0, k0t = 0.0;
      for (int slice = Q.slice1; slice <= Q.slice2; slice += skip)
        fsum += (slice == Q.slice1 || slice == Q.slice2) ? 0.5 : 1.0;
      lr_const_terms(R, S, Q.level, fsum, homo, background, k0t);
      U = ksum + homo;
      U += R.use_background ? background : k0t;
    }
    partials[(size_t)r * kblocks + blockIdx.x] = U;
  }
}

// out[r] = the request's block sums in block order.
__global__ void k_lr_finish(const double* __restrict__ partials, int kblocks, int nreq, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nreq) return;
  double t = partials[(size_t)r * kblocks];
  for (int b = 1; b < kblocks; b++) t += partials[(size_t)r * kblocks + b];
  out[r] = t;
}

int long_range_kblocks(int max_nk) { return (max_nk + kLrChunk - 1) / kLrChunk; }

bool long_range_phase_fits(int stride, int nptcls) {
  return (size_t)2 * nptcls * stride * sizeof(d2v) <= kLrPhaseLds;
}

template <bool MIXED, bool UPD>
static void launch_lr_group(int nsp, const ReplicaDev* reps, const ReqDev* reqs, int r0, int n, int kblocks,
                            size_t lds, double* partials, hipStream_t s) {
  const dim3 grid(kblocks, n);
  if (nsp <= 1)
    k_lr_req<MIXED, 1, UPD><<<grid, kBlock, lds, s>>>(reps, reqs, r0, kblocks, partials);
  else if (nsp == 2)
    k_lr_req<MIXED, 2, UPD><<<grid, kBlock, lds, s>>>(reps, reqs, r0, kblocks, partials);
  else
    k_lr_req<MIXED, kMaxSpecies, UPD><<<grid, kBlock, lds, s>>>(reps, reqs, r0, kblocks, partials);
}

std::vector<LrGroup> long_range_groups(int nreq, const int* replicas, const int* nspecies, const int* modes) {
  // consecutive requests of distinct replicas share a launch; a request on a replica already in
  // the launch starts the next one, so requests on one replica run in order
  std::vector<LrGroup> groups;
  std::unordered_map<int, int> last;   // replica -> the group that last holds it
  int r0 = 0;
  while (r0 < nreq) {
    LrGroup g{r0, r0, 0, false};
    const int gi = (int)groups.size();
    while (g.r1 < nreq && g.r1 - r0 < 65535) {
      auto it = last.find(replicas[g.r1]);
      if (it != last.end() && it->second == gi) break;
      last[replicas[g.r1]] = gi;
      g.nsp = std::max(g.nsp, nspecies[g.r1]);
      g.upd = g.upd || modes[g.r1] == 1;
      g.r1++;
    }
    groups.push_back(g);
    r0 = g.r1;
  }
  return groups;
}

void launch_long_range(bool mixed, const ReplicaDev* reps, const ReqDev* reqs, int nreq,
                       const std::vector<LrGroup>& groups, int kblocks, double* partials, double* out,
                       hipStream_t s) {
  for (const LrGroup& g : groups) {
    // phase rows only where a request updates rho_k NEW
    const size_t lds = kLrScratch + (g.upd ? kLrPhaseLds : 0);
    const int n = g.r1 - g.r0;
    if (mixed && g.upd)
      launch_lr_group<true, true>(g.nsp, reps, reqs, g.r0, n, kblocks, lds, partials, s);
    else if (g.upd)
      launch_lr_group<false, true>(g.nsp, reps, reqs, g.r0, n, kblocks, lds, partials, s);
    else if (mixed)
      launch_lr_group<true, false>(g.nsp, reps, reqs, g.r0, n, kblocks, lds, partials, s);
    else
      launch_lr_group<false, false>(g.nsp, reps, reqs, g.r0, n, kblocks, lds, partials, s);
  }
  k_lr_finish<<<(nreq + 255) / 256, 256, 0, s>>>(partials, kblocks, nreq, out);
}

}  // namespace pimc
