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

// Sum over this thread's k in [k0, k1) of the k-dependent terms of LongRangeClass::SingleAction
// on copy `mode` at one slice (homologous |rho|^2 u_k and heterologous 2 Re(rho_a rho_b*) u_k).
__device__ inline double lr_k_terms(const RepCore& R, int mode, int slice, int level, double factor, int k0, int k1) {
  double part = 0.0;
  const d2v* rho = R.rho(mode) + (size_t)slice * R.ns * R.nk;
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    for (int s = 0; s < R.ns; s++) {
      if (!((R.lr_mask >> (s * R.ns + s)) & 1u)) continue;
      const d2v r = rho[(size_t)s * R.nk + k];
      const double rhok2 = r.x * r.x + r.y * r.y;
      part += factor * 0.5 * 2.0 * rhok2 * R.ulong_k[((size_t)(s * R.ns + s) * R.lr_levels + level) * R.nk + k];
    }
    for (int s1 = 0; s1 < R.ns; s1++)
      for (int s2 = s1 + 1; s2 < R.ns; s2++) {
        if (!((R.lr_mask >> (s1 * R.ns + s2)) & 1u)) continue;
        const d2v a = rho[(size_t)s1 * R.nk + k], b = rho[(size_t)s2 * R.nk + k];
        const double rhorho = a.x * b.x + a.y * b.y;
        part += factor * 2.0 * rhorho * R.ulong_k[((size_t)(s1 * R.ns + s2) * R.lr_levels + level) * R.nk + k];
      }
  }
  return part;
}

// LongRangeClass::SingleAction for request req0 + blockIdx.y, over the k-vectors
// [blockIdx.x kLrChunk, (blockIdx.x + 1) kLrChunk) of its replica: in NEWMODE the block first
// brings those k of rho_k NEW up to date (NEW <- OLD over the range, then the moved particles'
// new phases added and old ones removed, PathClass::UpdateRho_ks(slice1, slice2, changed,
// level)).  partials[request][blockIdx.x] holds the block's sum; block 0 adds the
// k-independent terms.  The requests of one launch belong to different replicas.
constexpr int kLrChunk = 512;

template <bool MIXED>
__global__ void __launch_bounds__(kBlock) k_lr_req(const ReplicaDev* __restrict__ reps, const ReqDev* __restrict__ reqs,
                                                   int req0, int kblocks, double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  typedef typename std::conditional<MIXED, float, double>::type PT;
  PIMC_LDS PhaseRow<PT>* Cn = (PIMC_LDS PhaseRow<PT>*)smem_raw;
  const int r = req0 + (int)blockIdx.y;
  const ReqDev Q = reqs[r];
  const PIMC_GLOBAL int* qp = (const PIMC_GLOBAL int*)reqs[r].ptcls;   // indexed at run time: not from Q
  const ReplicaDev* RD = reps + Q.replica;
  const RepCore R = load_uniform(&RD->core);
  const RepSpecies& S = RD->sp;
  const int stride = kstride(R);
  PIMC_LDS PhaseRow<PT>* Co = Cn + stride;
  PIMC_LDS double* scratch = (PIMC_LDS double*)(Co + stride);   // 8-byte aligned for either entry size
  const int o1 = koff_of(R, 1), o2 = koff_of(R, 2);
  const int skip = 1 << Q.level;
  const int k0 = min(R.nk, (int)blockIdx.x * kLrChunk), k1 = min(R.nk, k0 + kLrChunk);

  if (Q.mode == 1) {
    // NEW <- OLD over the range, then add the moved particles' new phases and remove the old ones
    for (int slice = Q.slice1; slice <= Q.slice2; slice += skip)
      for (int sp = 0; sp < R.ns; sp++)
        for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
          const size_t at = ((size_t)slice * R.ns + sp) * R.nk + k;
          R.rho1[at] = R.rho0[at];
        }
    for (int pi = 0; pi < Q.nptcls; pi++) {
      const int p = qp[pi];
      int sp = 0;
      while (sp + 1 < R.ns && p >= S.first[sp + 1]) sp++;
      for (int slice = Q.slice1; slice <= Q.slice2; slice += skip) {
        __syncthreads();
        if (threadIdx.x < 6) {
          const int mode = threadIdx.x < 3 ? 1 : 0;
          const int d = threadIdx.x % 3;
          double x[3];
          load_bead(R, mode, slice, p, x);
          phase_rows_lane<PT>(R, x, d, mode ? Cn : Co);
        }
        __syncthreads();
        d2v* rho = R.rho1 + ((size_t)slice * R.ns + sp) * R.nk;
        for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
          const int ki[3] = {R.kidx[4 * k], R.kidx[4 * k + 1], R.kidx[4 * k + 2]};
          PT nr, ni, orr, oi;
          kterm<PT>(Cn, o1, o2, ki, nr, ni);
          kterm<PT>(Co, o1, o2, ki, orr, oi);
          d2v v = rho[k];
          v.x += (double)nr;
          v.y += (double)ni;
          v.x -= (double)orr;
          v.y -= (double)oi;
          rho[k] = v;
        }
      }
    }
    __syncthreads();
  }
  double part = 0.0;
  for (int slice = Q.slice1; slice <= Q.slice2; slice += skip) {
    const double factor = (slice == Q.slice1 || slice == Q.slice2) ? 0.5 : 1.0;
    part += lr_k_terms(R, Q.mode, slice, Q.level, factor, k0, k1);
  }
  const double ksum = block_sum(part, scratch);
  if (threadIdx.x == 0) {
    double U = ksum;
    if (blockIdx.x == 0) {
      double homo = 0.0, background = 0.0, k0t = 0.0;
      for (int slice = Q.slice1; slice <= Q.slice2; slice += skip) {
        const double factor = (slice == Q.slice1 || slice == Q.slice2) ? 0.5 : 1.0;
        lr_const_terms(R, S, Q.level, factor, homo, background, k0t);
      }
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

void launch_long_range(bool mixed, const ReplicaDev* reps, const ReqDev* reqs, int nreq, const int* strides,
                       const int* replicas, int kblocks, double* partials, double* out, hipStream_t s) {
  // consecutive requests of distinct replicas run in one launch; a request on a replica already
  // in the launch starts the next one, so requests on one replica run in order
  int r0 = 0;
  while (r0 < nreq) {
    int r1 = r0 + 1, stride = strides[r0];
    while (r1 < nreq && r1 - r0 < 65535) {
      bool seen = false;
      for (int q = r0; q < r1; q++) seen = seen || replicas[q] == replicas[r1];
      if (seen) break;
      stride = std::max(stride, strides[r1]);
      r1++;
    }
    const size_t lds = (size_t)2 * stride * 16 + kWaves * 8;
    const dim3 grid(kblocks, r1 - r0);
    if (mixed)
      k_lr_req<true><<<grid, kBlock, lds, s>>>(reps, reqs, r0, kblocks, partials);
    else
      k_lr_req<false><<<grid, kBlock, lds, s>>>(reps, reqs, r0, kblocks, partials);
    r0 = r1;
  }
  k_lr_finish<<<(nreq + 255) / 256, 256, 0, s>>>(partials, kblocks, nreq, out);
}

}  // namespace pimc
