// pimc_abi.hip -- host implementation of the C ABI in include/pimc_gpu.h.
//
// Owns device memory, builds the spline tables the kernels read, keeps the
// device copies of the fit and replica descriptors, and launches the kernels
// of pimc_actions.hip and pimc_bisection.hip on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <limits>
#include <vector>

#include "../../include/pimc_gpu.h"
#include "pimc_device.hpp"
#include "pimc_kernels.hpp"
#include "pimc_launch.hpp"

using namespace pimc;

namespace {

thread_local std::string g_err;

// Device memory of pimc_action_short_range_noalloc: request staging, per-tile partial
// sums and outputs.  One copy per GPU, at its own address on each; host code uses the
// address hipGetSymbolAddress returns for the current device (g_noalloc_ptr), never
// this name.
static __device__ __align__(256) unsigned char scratch[PIMC_NOALLOC_SCRATCH_BYTES];
constexpr int kNoallocDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoallocDevices] = {};

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(PIMC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device buffer owned by the context.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
  int ensure(size_t n) {
    if (n <= bytes) return PIMC_OK;
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) return fail(PIMC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    bytes = n;
    return PIMC_OK;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

// --------------------------------------------------------------------------
// Host spline setup: grids and MultiCubicSpline::Update (CubicSpline.cc:77-128)
// --------------------------------------------------------------------------
struct HostGrid {
  int type = 0, n = 0;
  double start = 0, end = 0, delta = 0, deltainv = 0, r0 = 0, spacing = 0;
  std::vector<double> x;
};

void make_linear(HostGrid& g, double s, double e, int n) {
  g.type = PIMC_GRID_LINEAR;
  g.start = s;
  g.end = e;
  g.n = n;
  g.delta = (e - s) / (double)(n - 1);
  g.deltainv = 1.0 / g.delta;
  g.x.resize(n);
  for (int i = 0; i < n; i++) g.x[i] = s + (double)i * g.delta;
}
// The .dm reader's LOG grid: spacing = (end/start)^(1/(n-1)), x_i = start * spacing^i.
void make_log(HostGrid& g, double s, double e, int n) {
  g.type = PIMC_GRID_LOG;
  g.spacing = pow((e / s), 1.0 / (n - 1.0));
  g.r0 = s;
  g.n = n;
  g.start = s;
  g.end = s * pow(g.spacing, (double)n - 1);
  g.x.resize(n);
  for (int i = 0; i < n; i++) g.x[i] = s * pow(g.spacing, (double)i);
}

// Solves the tridiagonal system lo[i] u[i-1] + di[i] u[i] + up[i] u[i+1] = rhs[i]
// (lo[0] and up[n-1] unused) by forward elimination and back substitution; the
// systems below are diagonally dominant, so no pivoting is needed.  rhs is overwritten
// with u.
void solve_tridiagonal(const std::vector<double>& lo, const std::vector<double>& di, const std::vector<double>& up,
                       std::vector<double>& rhs) {
  const size_t n = di.size();
  std::vector<double> upn(n);   // the super-diagonal after elimination, with a unit diagonal
  double piv = di[0];
  upn[0] = up[0] / piv;
  rhs[0] /= piv;
  for (size_t i = 1; i < n; i++) {
    piv = di[i] - lo[i] * upn[i - 1];
    upn[i] = i + 1 < n ? up[i] / piv : 0.0;
    rhs[i] = (rhs[i] - lo[i] * rhs[i - 1]) / piv;
  }
  for (size_t i = n - 1; i-- > 0;) rhs[i] -= upn[i] * rhs[i + 1];
}

// Knot second derivatives m_j of the C2 cubic spline through (x_j, y_j), y_j = y[j * stride]
// (the setup of MultiCubicSpline::Update, CubicSpline.cc:77-128).  With h_j = x_{j+1} - x_j
// and secant slopes s_j = (y_{j+1} - y_j) / h_j, continuity of the first derivative at an
// interior knot reads
//   h_{j-1} m_{j-1} + 2 (h_{j-1} + h_j) m_j + h_j m_{j+1} = 6 (s_j - s_{j-1}).
// An end with a given slope f adds 2 h_0 m_0 + h_0 m_1 = 6 (s_0 - f) (first knot) or
// h m_{n-2} + 2 h m_{n-1} = 6 (f - s_{n-2}) (last knot); a slope > 0.99e30 means a
// natural end, m = 0.
void spline_d2(const std::vector<double>& x, const double* y, int stride, double start_deriv, double end_deriv,
               std::vector<double>& d2) {
  const int n = (int)x.size();
  std::vector<double> lo(n, 0.0), di(n, 1.0), up(n, 0.0);
  d2.assign(n, 0.0);
  auto secant = [&](int j) { return (y[(size_t)(j + 1) * stride] - y[(size_t)j * stride]) / (x[j + 1] - x[j]); };
  for (int j = 1; j < n - 1; j++) {
    lo[j] = x[j] - x[j - 1];
    up[j] = x[j + 1] - x[j];
    di[j] = 2.0 * (lo[j] + up[j]);
    d2[j] = 6.0 * (secant(j) - secant(j - 1));
  }
  if (start_deriv <= 0.99e30) {
    const double h = x[1] - x[0];
    di[0] = 2.0 * h;
    up[0] = h;
    d2[0] = 6.0 * (secant(0) - start_deriv);
  }
  if (end_deriv <= 0.99e30) {
    const double h = x[n - 1] - x[n - 2];
    lo[n - 1] = h;
    di[n - 1] = 2.0 * h;
    d2[n - 1] = 6.0 * (end_deriv - secant(n - 2));
  }
  solve_tridiagonal(lo, di, up, d2);
}

struct HostFit {
  FitDev dev{};            // pointers refer to `mem` (and `dmem` for the beta-derivative tables)
  DevBuf mem, dmem;
  int num_terms = 0;       // -1 for fits without the DavidPA expansion
  size_t staged64 = 0, staged32 = 0;          // LDS bytes of one staged level
  size_t staged64_lin = 0, staged32_lin = 0;  // the same for the linear-grid kernels (coefficients only)
  int num_levels() const { return dev.num_tau - dev.taupos; }
};

// Knot slopes d_j of the C2 cubic through the values v_j = v[j * stride] along one line
// of a bicubic table, written to d[j * stride] (the slopes BicubicSpline::XUpdate /
// YUpdate / BiUpdate set up, BicubicSpline.cc:20-154).  On [x_j, x_{j+1}] the cubic
// Hermite piece with end slopes d_j, d_{j+1} has second derivative
// (6 s_j - 4 d_j - 2 d_{j+1}) / h_j at its left end and (4 d_{j+1} + 2 d_j - 6 s_j) / h_j
// at its right end (h_j = x_{j+1} - x_j, s_j the secant slope).  Equal second
// derivatives at an interior knot give
//   h_j d_{j-1} + 2 (h_{j-1} + h_j) d_j + h_{j-1} d_{j+1} = 3 (h_j s_{j-1} + h_{j-1} s_j),
// and zero second derivatives at the ends (natural) give 2 d_0 + d_1 = 3 s_0 and
// d_{n-2} + 2 d_{n-1} = 3 s_{n-2}.
void bc_slopes(const std::vector<double>& x, const double* v, double* d, size_t stride) {
  const int n = (int)x.size();
  std::vector<double> lo(n, 0.0), di(n, 2.0), up(n, 0.0), rhs(n);
  auto secant = [&](int j) { return (v[(size_t)(j + 1) * stride] - v[(size_t)j * stride]) / (x[j + 1] - x[j]); };
  up[0] = 1.0;
  rhs[0] = 3.0 * secant(0);
  for (int j = 1; j < n - 1; j++) {
    const double hl = x[j] - x[j - 1], hr = x[j + 1] - x[j];
    lo[j] = hr;
    di[j] = 2.0 * (hl + hr);
    up[j] = hl;
    rhs[j] = 3.0 * (hr * secant(j - 1) + hl * secant(j));
  }
  lo[n - 1] = 1.0;
  rhs[n - 1] = 3.0 * secant(n - 2);
  solve_tridiagonal(lo, di, up, rhs);
  for (int j = 0; j < n; j++) d[(size_t)j * stride] = rhs[j];
}

int make_host_grid(int type, double start, double end, int n, HostGrid& g) {
  if (n < 3 || !(start < end)) return fail(PIMC_ERR_ARG, "grid: bad range or size");
  if (type == PIMC_GRID_LINEAR)
    make_linear(g, start, end, n);
  else if (type == PIMC_GRID_LOG && start > 0)
    make_log(g, start, end, n);
  else
    return fail(PIMC_ERR_ARG, "grid: bad type");
  return PIMC_OK;
}

struct HostReplica {
  ReplicaDev dev{};
  DevBuf path0, path1, perm0, perm1;
  DevBuf joinbuf;                // the replica's join (one int, device-resident: the chain moves it)
  DevBuf rho0, rho1, kidx_dev, ulong_k, lr_const, ulong, ulong32;
  DevBuf dulong, dulong_k, dlr_const;   // beta derivatives (energy estimator)
  DevBuf vlong, vlong_k, vlr_const;     // long-range potential (potential-energy estimator)
  unsigned dulong_mask = 0, dlr_mask = 0;
  std::vector<double> kvecs;     // [nk][3]
  std::vector<int> kidx;         // [nk][3] index + maxk
  bool use_long_range = false;
  int M = 0;
  // set once a cycle has been applied: the permutation may no longer be the identity
  bool maybe_permuted = false;
  // set by the resident chain, which updates rho_k OLD only: copy NEW is refreshed from OLD
  // (refresh_rho1) before any call that reads or writes it
  bool rho1_stale = false;
  // PermuteType TABLE for moves of permute_species (pimc_replica_set_permute)
  bool permute_on = false;
  int permute_species = 0;
  PermuteDev permute{};
};

// PathClass::SetupkVecs3D with Include (PathClass.cc:281-292, 534-586): integer
// triples inside |k| < kc in the half space, in ix, iy, iz loop order.
void setup_kvecs(HostReplica& H, double kcut) {
  RepCore& R = H.dev.core;
  int maxk[3];
  for (int i = 0; i < 3; i++) {
    R.kbox[i] = 2.0 * M_PI / R.box[i];
    maxk[i] = (int)ceil(1.1 * kcut / R.kbox[i]);
  }
  R.maxk0 = maxk[0];
  R.maxk1 = maxk[1];
  R.maxk2 = maxk[2];
  H.kvecs.clear();
  H.kidx.clear();
  for (int ix = -maxk[0]; ix <= maxk[0]; ix++) {
    double k0 = ix * R.kbox[0];
    for (int iy = -maxk[1]; iy <= maxk[1]; iy++) {
      double k1 = iy * R.kbox[1];
      for (int iz = -maxk[2]; iz <= maxk[2]; iz++) {
        double k2 = iz * R.kbox[2];
        bool include = (k0 > 0.0) || ((k0 == 0.0) && (k1 > 0.0)) || ((k0 == 0.0) && (k1 == 0.0) && (k2 > 0.0));
        if ((k0 * k0 + k1 * k1 + k2 * k2 < kcut * kcut) && include) {
          H.kvecs.insert(H.kvecs.end(), {k0, k1, k2});
          H.kidx.insert(H.kidx.end(), {ix + maxk[0], iy + maxk[1], iz + maxk[2]});
        }
      }
    }
  }
  R.nk = (int)(H.kvecs.size() / 3);
}

}  // namespace

// A captured sweep and the device state its replays use.
struct SweepGraph {
  DevBuf ids, pos, stats, queue, first, bar, part;
  std::vector<int> reps;   // the replicas a replay moves
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int num_moves = 0;
  int64_t first_move = 0, sweeps = 0;
  SweepGraph() = default;
  SweepGraph(const SweepGraph&) = delete;
  SweepGraph& operator=(const SweepGraph&) = delete;
  ~SweepGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct pimc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<HostFit>> fits;
  std::vector<std::unique_ptr<HostReplica>> reps;
  DevBuf d_fits, d_reps;
  bool fits_dirty = true, reps_dirty = true;
  DevBuf d_reqs, d_partials, d_out, d_in0, d_in1, d_in2, d_idx, d_img, d_records, d_stats, d_repids;
  // Pinned, coherent host memory for the requests and results of small action calls (up to
  // kPinnedReqs requests): the kernels read the requests and write the results there
  // directly, so a call is its launches and one synchronisation, with no copies.  Allocated
  // on first use, freed with the context.
  void* h_pin = nullptr;
  void* d_pin = nullptr;
  bool pin_busy = false;   // a call staged its requests there and has not synchronised yet
  // bisection counters accumulated per replica id ([acc_cap][5]) until pimc_bisection_stats_take
  DevBuf d_acc;
  int acc_cap = 0;
  std::vector<int> h_repids, h_reppos;       // launch list of the last upload to d_repids / d_reppos
  DevBuf d_reppos, d_queue;
  // cluster launches: barrier counters, partial sums, the sticky abort flag, and
  // the bound of one barrier wait in device wall-clock ticks
  DevBuf d_bar, d_part, d_abort;
  DevBuf d_permute;   // per list entry of a TABLE launch: its cycle weights and cut-off
  DevBuf d_cycles;    // cycle-table scratch of TABLE launches
  uint64_t wait_ticks = 0;
  pimc_bisection_plan plan{256, 1, 0, 1, 0.0};   // launch shape of pimc_bisection_run
  std::vector<std::unique_ptr<SweepGraph>> graphs;
  // event pairs around timed launches since the last take, and a pool of spare ones; past
  // kMaxTimed pairs the finished ones are folded into timed_ms
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed, spare;
  double timed_ms = 0.0;

  // Host staging of the request upload; kept alive until the call that uses it synchronizes.
  std::vector<ReqDev> h_reqs;

  // A live sweep graph holds device addresses of the descriptor arrays and counters.
  bool graphs_alive() const {
    for (auto& g : graphs)
      if (g) return true;
    return false;
  }
  int sync_descriptors() {
    std::vector<FitDev> hf;
    std::vector<ReplicaDev> hr;
    if (!(fits_dirty && !fits.empty()) && !(reps_dirty && !reps.empty())) return PIMC_OK;
    if (graphs_alive() && (fits.size() * sizeof(FitDev) > d_fits.bytes || reps.size() * sizeof(ReplicaDev) > d_reps.bytes))
      return fail(PIMC_ERR_STATE, "fits or replicas were added while a sweep graph exists; destroy the graphs first");
    if (fits_dirty && !fits.empty()) {
      for (auto& f : fits) hf.push_back(f->dev);
      int rc = d_fits.ensure(hf.size() * sizeof(FitDev));
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(d_fits.p, hf.data(), hf.size() * sizeof(FitDev), hipMemcpyHostToDevice, stream));
      fits_dirty = false;
    }
    if (reps_dirty && !reps.empty()) {
      for (auto& r : reps) hr.push_back(r->dev);
      int rc = d_reps.ensure(hr.size() * sizeof(ReplicaDev));
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(d_reps.p, hr.data(), hr.size() * sizeof(ReplicaDev), hipMemcpyHostToDevice, stream));
      reps_dirty = false;
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return PIMC_OK;
  }
  HostReplica* rep(int id) { return (id >= 0 && id < (int)reps.size()) ? reps[id].get() : nullptr; }
};

namespace {

// An event pair for a timed launch (bisection runs, sweep replays, long-range action calls),
// queued for pimc_bisection_stats_take.  The list stays bounded for callers that never take:
// past kMaxTimed pairs the stream is drained and the pairs are folded into ctx->timed_ms.
constexpr size_t kMaxTimed = 1024;
int timed_pair(pimc_ctx* ctx, std::pair<hipEvent_t, hipEvent_t>& ev) {
  if (ctx->timed.size() >= kMaxTimed) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto& e : ctx->timed) {
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
      ctx->timed_ms += ms;
      ctx->spare.push_back(e);
    }
    ctx->timed.clear();
  }
  if (!ctx->spare.empty()) {
    ev = ctx->spare.back();
    ctx->spare.pop_back();
  } else {
    HIP_TRY(hipEventCreate(&ev.first));
    HIP_TRY(hipEventCreate(&ev.second));
  }
  ctx->timed.push_back(ev);
  return PIMC_OK;
}

// Copy NEW of rho_k <- copy OLD where a chain run left NEW behind (stream-ordered after it).
int refresh_rho1(pimc_ctx* ctx, HostReplica* H) {
  if (!H->rho1_stale) return PIMC_OK;
  const RepCore& R = H->dev.core;
  HIP_TRY(hipMemcpyAsync(H->rho1.p, H->rho0.p, (size_t)R.nslices * R.ns * R.nk * sizeof(d2v),
                         hipMemcpyDeviceToDevice, ctx->stream));
  H->rho1_stale = false;
  return PIMC_OK;
}


int check_req(pimc_ctx* ctx, const pimc_action_req& q) {
  HostReplica* H = ctx->rep(q.replica);
  if (!H) return fail(PIMC_ERR_ARG, "action request: bad replica id");
  if (q.mode != PIMC_OLDMODE && q.mode != PIMC_NEWMODE) return fail(PIMC_ERR_ARG, "action request: bad mode");
  if (q.level < 0 || q.level >= PIMC_MAX_LEVELS) return fail(PIMC_ERR_ARG, "action request: bad level");
  if (q.slice1 < 0 || q.slice2 > H->M || q.slice1 > q.slice2)
    return fail(PIMC_ERR_ARG, "action request: bad slice range");
  // the link loop reads slice + 2^level for every slice < slice2
  const int skip = 1 << q.level;
  if (q.slice2 > q.slice1 && q.slice1 + ((q.slice2 - q.slice1 + skip - 1) / skip) * skip > H->M)
    return fail(PIMC_ERR_ARG, "action request: slice range runs past the last slice at this level");
  if (q.nptcls < 1 || q.nptcls > PIMC_MAX_ACTIVE) return fail(PIMC_ERR_ARG, "action request: bad particle count");
  for (int i = 0; i < q.nptcls; i++)
    if (q.ptcls[i] < 0 || q.ptcls[i] >= H->dev.core.N) return fail(PIMC_ERR_ARG, "action request: bad particle index");
  return PIMC_OK;
}

// Requests of one call and where its results go: small calls use the context's pinned host
// region (requests read, results written by the kernels in place), larger ones device buffers
// with a copy each way.
constexpr int kPinnedReqs = 64;
constexpr size_t kPinnedOutOffset = 4096;
static_assert(kPinnedReqs * sizeof(ReqDev) <= kPinnedOutOffset, "pinned request region");
constexpr size_t kPinnedBytes = kPinnedOutOffset + kPinnedReqs * sizeof(double);
struct CallIO {
  const ReqDev* reqs = nullptr;   // device-visible
  double* out = nullptr;          // device-visible
  bool pinned = false;
};

int upload_reqs(pimc_ctx* ctx, const pimc_action_req* reqs, int nreq);

int stage_call(pimc_ctx* ctx, const pimc_action_req* reqs, int nreq, CallIO& io) {
  if (nreq > kPinnedReqs) {
    int rc = upload_reqs(ctx, reqs, nreq);
    if (!rc) rc = ctx->d_out.ensure((size_t)nreq * sizeof(double));
    if (rc) return rc;
    io.reqs = ctx->d_reqs.as<ReqDev>();
    io.out = ctx->d_out.as<double>();
    io.pinned = false;
    return PIMC_OK;
  }
  for (int i = 0; i < nreq; i++)
    if (int rc = check_req(ctx, reqs[i])) return rc;
  if (!ctx->h_pin) {
    HIP_TRY(hipHostMalloc(&ctx->h_pin, kPinnedBytes, hipHostMallocMapped | hipHostMallocCoherent));
    if (hipHostGetDevicePointer(&ctx->d_pin, ctx->h_pin, 0) != hipSuccess) {
      (void)hipHostFree(ctx->h_pin);
      ctx->h_pin = nullptr;
      return fail(PIMC_ERR_DEVICE, "action call: no device address for the pinned request region");
    }
  }
  // a call that returned an error after staging may have left kernels that still read the region
  if (ctx->pin_busy) HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->pin_busy = true;
  ReqDev* h = (ReqDev*)ctx->h_pin;
  for (int i = 0; i < nreq; i++) {
    ReqDev q{};
    q.replica = reqs[i].replica;
    q.mode = reqs[i].mode;
    q.slice1 = reqs[i].slice1;
    q.slice2 = reqs[i].slice2;
    q.level = reqs[i].level;
    q.nptcls = reqs[i].nptcls;
    for (int k = 0; k < PIMC_MAX_ACTIVE; k++) q.ptcls[k] = k < reqs[i].nptcls ? reqs[i].ptcls[k] : -1;
    h[i] = q;
  }
  io.reqs = (const ReqDev*)ctx->d_pin;
  io.out = (double*)((unsigned char*)ctx->d_pin + kPinnedOutOffset);
  io.pinned = true;
  return PIMC_OK;
}

// After the call's launches: the results into out, and the one synchronisation.
int finish_call(pimc_ctx* ctx, const CallIO& io, int nreq, double* out) {
  if (!io.pinned)
    HIP_TRY(hipMemcpyAsync(out, io.out, (size_t)nreq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (io.pinned) {
    memcpy(out, (const unsigned char*)ctx->h_pin + kPinnedOutOffset, (size_t)nreq * sizeof(double));
    ctx->pin_busy = false;
  }
  return PIMC_OK;
}

// Link ranges per partner tile of a short-range call (LaunchShortRange::nsplit): a call with
// few requests splits every tile's links over up to 8 blocks, so that each thread walks a few
// links rather than the whole window.  Both short-range entry points use it, so they give the
// same sums bit for bit.
int short_range_split(int nreq, int ntiles) { return nreq * ntiles >= 256 ? 1 : std::min(8, 256 / (nreq * ntiles)); }

int upload_reqs(pimc_ctx* ctx, const pimc_action_req* reqs, int nreq) {
  std::vector<ReqDev>& h = ctx->h_reqs;
  h.assign(nreq, ReqDev{});
  for (int i = 0; i < nreq; i++) {
    int rc = check_req(ctx, reqs[i]);
    if (rc) return rc;
    h[i].replica = reqs[i].replica;
    h[i].mode = reqs[i].mode;
    h[i].slice1 = reqs[i].slice1;
    h[i].slice2 = reqs[i].slice2;
    h[i].level = reqs[i].level;
    h[i].nptcls = reqs[i].nptcls;
    for (int k = 0; k < PIMC_MAX_ACTIVE; k++) h[i].ptcls[k] = k < reqs[i].nptcls ? reqs[i].ptcls[k] : -1;
  }
  int rc = ctx->d_reqs.ensure(h.size() * sizeof(ReqDev));
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->d_reqs.p, h.data(), h.size() * sizeof(ReqDev), hipMemcpyHostToDevice, ctx->stream));
  return PIMC_OK;
}

// Host [slice][ptcl][3] <-> device [slice][3][N]
void to_device_layout(const double* src, double* dst, int nslices, int N) {
  for (int s = 0; s < nslices; s++)
    for (int p = 0; p < N; p++)
      for (int d = 0; d < 3; d++) dst[((size_t)s * 3 + d) * N + p] = src[((size_t)s * N + p) * 3 + d];
}
void to_host_layout(const double* src, double* dst, int nslices, int N) {
  for (int s = 0; s < nslices; s++)
    for (int p = 0; p < N; p++)
      for (int d = 0; d < 3; d++) dst[((size_t)s * N + p) * 3 + d] = src[((size_t)s * 3 + d) * N + p];
}

}  // namespace

extern "C" {

int pimc_abi_version(void) { return PIMC_ABI_VERSION; }
const char* pimc_last_error(void) { return g_err.c_str(); }

int pimc_ctx_create(int32_t device, pimc_ctx** out) {
  if (!out) return fail(PIMC_ERR_ARG, "ctx_create: out is NULL");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(PIMC_ERR_DEVICE, std::string("ctx_create: no GPU device (") + hipGetErrorString(e) + ")");
  if (device < 0 || device >= count) return fail(PIMC_ERR_ARG, "ctx_create: bad device index");
  HIP_TRY(hipSetDevice(device));
  if (device < kNoallocDevices) {
    // once per device, with the device current: the device address of its scratch array
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
      void* ptr = nullptr;
      HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      if (!ptr) return fail(PIMC_ERR_DEVICE, "ctx_create: no device address for the scratch array");
      g_noalloc_ptr[device] = ptr;
    }
  }
  auto c = std::make_unique<pimc_ctx>();
  c->device = device;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
    khz = 100000;   // the MI300-series wall clock runs at 100 MHz
  // bound of one cluster barrier wait: two seconds, or PIMC_CLUSTER_WAIT_US microseconds
  // (a test can make every wait give up at once to exercise the abort path)
  double wait_us = 2.0e6;
  if (const char* env = getenv("PIMC_CLUSTER_WAIT_US")) wait_us = std::max(0.0, atof(env));
  c->wait_ticks = (uint64_t)((double)khz * 1.0e-3 * wait_us);
  int rc = c->d_abort.ensure(sizeof(int));
  if (rc) return rc;
  HIP_TRY(hipMemset(c->d_abort.p, 0, sizeof(int)));
  *out = c.release();
  return PIMC_OK;
}

int pimc_ctx_destroy(pimc_ctx* ctx) {
  if (!ctx) return PIMC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  hipStream_t s = ctx->stream;
  for (auto* v : {&ctx->timed, &ctx->spare})
    for (auto& e : *v) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
  return PIMC_OK;
}

int pimc_ctx_synchronize(pimc_ctx* ctx) {
  if (!ctx) return fail(PIMC_ERR_ARG, "NULL ctx");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

// Interval cubics {y_lo, ..} in the local coordinate of one table column (see FitDev),
// from the column's natural-start, zero-end-slope spline (DavidPAClass.cc:760-763).
void column_cubics(const HostGrid& g, const double* y, int stride, std::vector<double>& d2, d4v* out, int out_stride) {
  spline_d2(g.x, y, stride, 5.0e30, 0.0, d2);
  for (int i = 0; i < g.n - 1; i++) {
    double h = g.x[i + 1] - g.x[i];
    double h2over6 = h * h * 0.1666666666666666666;
    double ylo = y[(size_t)i * stride], yhi = y[(size_t)(i + 1) * stride];
    double Dlo = d2[i] * h2over6, Dhi = d2[i + 1] * h2over6;
    out[(size_t)i * out_stride] = d4v{ylo, yhi - ylo - 2.0 * Dlo - Dhi, 3.0 * Dlo, Dhi - Dlo};
  }
}

int pimc_fit_create_david(pimc_ctx* ctx, const pimc_david_fit_desc* d, int32_t* fit_id) {
  if (!ctx || !d || !fit_id || !d->y) return fail(PIMC_ERR_ARG, "fit_create_david: NULL argument");
  if (d->num_points < 3 || d->num_tau < 1 || d->num_cols < 2)
    return fail(PIMC_ERR_ARG, "fit_create_david: bad sizes");
  const int nt = d->num_terms;
  if (nt < 0 || nt > 3) return fail(PIMC_ERR_UNSUPPORTED, "fit_create_david: num_terms must be 0..3");
  const int noff = nt * (nt + 3) / 2;
  if (d->num_cols < 2 + noff) return fail(PIMC_ERR_ARG, "fit_create_david: num_cols too small for num_terms");
  if (!(d->grid_start < d->grid_end)) return fail(PIMC_ERR_ARG, "fit_create_david: bad grid range");
  HostGrid g;
  if (d->grid_type == PIMC_GRID_LINEAR)
    make_linear(g, d->grid_start, d->grid_end, d->num_points);
  else if (d->grid_type == PIMC_GRID_LOG && d->grid_start > 0)
    make_log(g, d->grid_start, d->grid_end, d->num_points);
  else
    return fail(PIMC_ERR_ARG, "fit_create_david: bad grid type");
  int taupos = -1;
  double tau = d->smallest_tau;
  for (int i = 0; i < d->num_tau; i++) {
    if (fabs(tau - d->desired_tau) < 1e-4) taupos = i;
    tau = tau * 2;
  }
  if (taupos < 0) return fail(PIMC_ERR_ARG, "fit_create_david: desired_tau matches no tau level");

  const int ng = g.n, nint = ng - 1, ncol = 1 + noff, ntau = d->num_tau, nc = d->num_cols;
  std::vector<d2v> geo(nint);
  for (int i = 0; i < nint; i++) {
    double h = g.x[i + 1] - g.x[i];
    geo[i] = d2v{g.x[i], 1.0 / h};
  }
  std::vector<d4v> c64((size_t)ntau * nint * ncol);
  std::vector<f4v> c32(c64.size());
  std::vector<unsigned char> big((size_t)ntau * nint, 0);
  std::vector<double> d2;
  for (int lev = 0; lev < ntau; lev++) {
    const double* y = d->y + (size_t)lev * ng * nc;
    for (int c = 0; c < ncol; c++) {
      const int col = 1 + c;  // table column: u00, then the off-diagonal columns
      // DavidPAClass builds its splines with natural start and zero end slope (DavidPAClass.cc:760-763)
      spline_d2(g.x, y + col, nc, 5.0e30, 0.0, d2);
      for (int i = 0; i < nint; i++) {
        double h = g.x[i + 1] - g.x[i];
        double h2over6 = h * h * 0.1666666666666666666;
        double ylo = y[(size_t)i * nc + col], yhi = y[(size_t)(i + 1) * nc + col];
        double Dlo = d2[i] * h2over6, Dhi = d2[i + 1] * h2over6;
        // a*ylo + b*yhi + (a^3-a)Dlo + (b^3-b)Dhi with a = 1-b, as a cubic in b
        d4v cf = d4v{ylo, yhi - ylo - 2.0 * Dlo - Dhi, 3.0 * Dlo, Dhi - Dlo};
        size_t at = ((size_t)lev * nint + i) * ncol + c;
        c64[at] = cf;
        c32[at] = f4v{(float)cf.x, (float)cf.y, (float)cf.z, (float)cf.w};
        if (fabs(cf.x) + fabs(cf.y) + fabs(cf.z) + fabs(cf.w) > kMixedLimit) big[(size_t)lev * nint + i] = 1;
      }
    }
    // FP32 rows of flagged intervals hold NaN: the mixed kernels read the FP64 row there
    for (int i = 0; i < nint; i++)
      if (big[(size_t)lev * nint + i])
        for (int c = 0; c < ncol; c++) {
          const float qnan = std::numeric_limits<float>::quiet_NaN();
          c32[((size_t)lev * nint + i) * ncol + c] = f4v{qnan, qnan, qnan, qnan};
        }
  }
  auto f = std::make_unique<HostFit>();
  size_t bx = (size_t)ng * 8, bgeo = (size_t)nint * 16, b64 = c64.size() * 32, b32 = c32.size() * 16;
  // per level: the first interval above every flagged one of that level
  std::vector<int> i32_lev(ntau, 0);
  for (int lev = 0; lev < ntau; lev++)
    for (int i = 0; i < nint; i++)
      if (big[(size_t)lev * nint + i]) i32_lev[lev] = i + 1;
  size_t off_geo = 0, off_x = off_geo + bgeo, off64 = (off_x + bx + 255) & ~(size_t)255,
         off32 = off64 + b64, off_i32 = (off32 + b32 + 15) & ~(size_t)15, total = off_i32 + (size_t)ntau * 4;
  int rc = f->mem.ensure(total);
  if (rc) return rc;
  unsigned char* base = f->mem.as<unsigned char>();
  HIP_TRY(hipMemcpy(base + off_geo, geo.data(), bgeo, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_x, g.x.data(), bx, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off64, c64.data(), b64, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off32, c32.data(), b32, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_i32, i32_lev.data(), (size_t)ntau * 4, hipMemcpyHostToDevice));
  if (d->dy) {
    // energy-estimator table: V (column 0 of y), du00 and the off-diagonal dukj columns
    const int ecol = ncol + 1;
    std::vector<d4v> e64((size_t)ntau * nint * ecol);
    for (int lev = 0; lev < ntau; lev++) {
      const double* y = d->y + (size_t)lev * ng * nc;
      const double* dy = d->dy + (size_t)lev * ng * nc;
      d4v* out = e64.data() + (size_t)lev * nint * ecol;
      column_cubics(g, y, nc, d2, out, ecol);
      for (int c = 1; c < ecol; c++) column_cubics(g, dy + c, nc, d2, out + c, ecol);
    }
    if ((rc = f->dmem.ensure(e64.size() * sizeof(d4v)))) return rc;
    HIP_TRY(hipMemcpy(f->dmem.p, e64.data(), e64.size() * sizeof(d4v), hipMemcpyHostToDevice));
    f->dev.e64 = f->dmem.as<d4v>();
    f->dev.ecol = ecol;
  }
  FitDev& F = f->dev;
  F.grid_type = g.type;
  F.ng = ng;
  F.nint = nint;
  F.num_tau = ntau;
  F.ncol = ncol;
  F.noff = noff;
  F.num_terms = nt;
  // the first interval above every flagged one (all levels)
  F.i32 = 0;
  for (int lev = 0; lev < ntau; lev++) F.i32 = std::max(F.i32, i32_lev[lev]);
  F.i32_lev = (const int*)(base + off_i32);
  F.taupos = taupos;
  F.start = g.start;
  F.end = g.end;
  F.deltainv = g.deltainv;
  F.r0 = g.r0;
  F.spacing = g.spacing;
  F.logspacing = g.type == PIMC_GRID_LOG ? log(g.spacing) : 0.0;
  F.x = (const double*)(base + off_x);
  F.geo = (const d2v*)(base + off_geo);
  F.c64 = (const d4v*)(base + off64);
  F.c32 = (const f4v*)(base + off32);
  f->num_terms = nt;
  f->staged64 = staged_bytes(nint, ncol, ng, false);
  F.i32_lds = (size_t)F.i32 * ncol * 32 <= (size_t)kFp64StageBytes ? F.i32 : 0;
  f->staged32 = staged_bytes(nint, ncol, ng, true, false, F.i32_lds);
  f->staged64_lin = staged_bytes(nint, ncol, ng, false, true);
  f->staged32_lin = staged_bytes(nint, ncol, ng, true, true, F.i32_lds);
  ctx->fits.push_back(std::move(f));
  ctx->fits_dirty = true;
  *fit_id = (int32_t)ctx->fits.size() - 1;
  return PIMC_OK;
}

int pimc_fit_create_coulomb_bc(pimc_ctx* ctx, const pimc_coulomb_bc_fit_desc* d, int32_t* fit_id) {
  if (!ctx || !d || !fit_id || !d->umat) return fail(PIMC_ERR_ARG, "fit_create_coulomb_bc: NULL argument");
  if (d->num_betas < 1 || d->num_betas > PIMC_MAX_LEVELS) return fail(PIMC_ERR_ARG, "fit_create_coulomb_bc: bad num_betas");
  HostGrid qg, tg;
  int rc = make_host_grid(d->q_grid_type, d->q_start, d->q_end, d->q_points, qg);
  if (rc) return rc;
  rc = make_host_grid(d->t_grid_type, d->t_start, d->t_end, d->t_points, tg);
  if (rc) return rc;
  const int nq = qg.n, nt = tg.n, nb = d->num_betas;
  // BCpoint {z, dzdx, dzdy, d2zdxdy} per node and level (BicubicSpline::Init)
  std::vector<double> z((size_t)nq * nt), dx(z.size()), dy(z.size()), dxy(z.size());
  std::vector<d4v> bc((size_t)nb * nq * nt);
  for (int b = 0; b < nb; b++) {
    const double* u = d->umat + (size_t)b * nq * nt;
    std::copy(u, u + z.size(), z.begin());
    for (int iy = 0; iy < nt; iy++) bc_slopes(qg.x, z.data() + iy, dx.data() + iy, nt);   // XUpdate
    for (int ix = 0; ix < nq; ix++) bc_slopes(tg.x, z.data() + (size_t)ix * nt, dy.data() + (size_t)ix * nt, 1);
    for (int iy = 0; iy < nt; iy++) bc_slopes(qg.x, dy.data() + iy, dxy.data() + iy, nt);  // BiUpdate
    for (size_t i = 0; i < z.size(); i++) bc[(size_t)b * z.size() + i] = d4v{z[i], dx[i], dy[i], dxy[i]};
  }
  auto f = std::make_unique<HostFit>();
  if (d->dumat) {
    // dU/dbeta bicubic: the same node-derivative construction on dumat
    std::vector<d4v> dbc(bc.size());
    for (int b = 0; b < nb; b++) {
      const double* u = d->dumat + (size_t)b * nq * nt;
      std::copy(u, u + z.size(), z.begin());
      for (int iy = 0; iy < nt; iy++) bc_slopes(qg.x, z.data() + iy, dx.data() + iy, nt);
      for (int ix = 0; ix < nq; ix++) bc_slopes(tg.x, z.data() + (size_t)ix * nt, dy.data() + (size_t)ix * nt, 1);
      for (int iy = 0; iy < nt; iy++) bc_slopes(qg.x, dy.data() + iy, dxy.data() + iy, nt);
      for (size_t i = 0; i < z.size(); i++) dbc[(size_t)b * z.size() + i] = d4v{z[i], dx[i], dy[i], dxy[i]};
    }
    if ((rc = f->dmem.ensure(dbc.size() * sizeof(d4v)))) return rc;
    HIP_TRY(hipMemcpy(f->dmem.p, dbc.data(), dbc.size() * sizeof(d4v), hipMemcpyHostToDevice));
    f->dev.dbc = f->dmem.as<d4v>();
  }
  // The mixed path: each node's derivatives {dz/dx, dz/dy, d2z/dxdy} in FP32 (it evaluates a
  // cell anchored on its first corner's FP64 value, so FP32 rounds the variation across the
  // cell, not U), and per level the first q interval from which every cell has sum over its
  // corners of |z - z_first| + h|dz/dx| + k|dz/dy| + hk|d2z/dxdy| within kBcMixedLimit: the
  // cells below it, deep in the core, are evaluated from the FP64 nodes.  The bound is loose
  // so that the lanes of a wave rarely split between the two evaluations.
  std::vector<f4v> bc32(bc.size());
  for (size_t i = 0; i < bc.size(); i++) bc32[i] = f4v{(float)bc[i].y, (float)bc[i].z, (float)bc[i].w, 0.0f};
  std::vector<int> i32lev(nb, 0);
  for (int b = 0; b < nb; b++) {
    int first = 0;
    for (int ix = 0; ix + 1 < nq; ix++) {
      const double h = qg.x[ix + 1] - qg.x[ix];
      bool ok = true;
      for (int iy = 0; iy + 1 < nt && ok; iy++) {
        const double k = tg.x[iy + 1] - tg.x[iy];
        double m = 0.0;
        const double z0 = bc[((size_t)b * nq + ix) * nt + iy].x;
        for (int cx = 0; cx < 2; cx++)
          for (int cy = 0; cy < 2; cy++) {
            const d4v& P = bc[((size_t)b * nq + ix + cx) * nt + iy + cy];
            m += fabs(P.x - z0) + h * fabs(P.y) + k * fabs(P.z) + h * k * fabs(P.w);
          }
        ok = m <= kBcMixedLimit;
      }
      if (!ok) first = ix + 1;
    }
    i32lev[b] = first;
  }
  // interval records {X_i, h_i, 1/h_i, X_{i+1}} of both grids
  auto records = [](const HostGrid& g) {
    std::vector<d4v> r(g.n - 1);
    for (int i = 0; i + 1 < g.n; i++) {
      const double h = g.x[i + 1] - g.x[i];
      r[i] = d4v{g.x[i], h, 1.0 / h, g.x[i + 1]};
    }
    return r;
  };
  const std::vector<d4v> qiv = records(qg), tiv = records(tg);
  size_t bq = (size_t)nq * 8, bt = (size_t)nt * 8, bb = bc.size() * sizeof(d4v);
  size_t bqi = qiv.size() * sizeof(d4v), bti = tiv.size() * sizeof(d4v);
  size_t off_q = 0, off_t = (bq + 255) & ~(size_t)255, off_b = (off_t + bt + 255) & ~(size_t)255;
  size_t off_qi = (off_b + bb + 255) & ~(size_t)255, off_ti = (off_qi + bqi + 255) & ~(size_t)255;
  size_t off_b32 = (off_ti + bti + 255) & ~(size_t)255, off_il = (off_b32 + bc32.size() * sizeof(f4v) + 255) & ~(size_t)255;
  size_t total = off_il + (size_t)nb * sizeof(int);
  if ((rc = f->mem.ensure(total))) return rc;
  unsigned char* base = f->mem.as<unsigned char>();
  HIP_TRY(hipMemcpy(base + off_q, qg.x.data(), bq, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_t, tg.x.data(), bt, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_b, bc.data(), bb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_qi, qiv.data(), bqi, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_ti, tiv.data(), bti, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_b32, bc32.data(), bc32.size() * sizeof(f4v), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(base + off_il, i32lev.data(), (size_t)nb * sizeof(int), hipMemcpyHostToDevice));
  FitDev& F = f->dev;
  F.kind = 1;
  F.c32 = (const f4v*)(base + off_b32);
  F.i32_lev = (const int*)(base + off_il);
  F.i32 = *std::max_element(i32lev.begin(), i32lev.end());
  F.grid_type = qg.type;
  F.ng = nq;
  F.nint = nq - 1;
  F.num_tau = nb;
  F.taupos = 0;
  F.start = qg.start;
  F.end = qg.end;
  F.deltainv = qg.deltainv;
  F.r0 = qg.r0;
  F.spacing = qg.spacing;
  F.logspacing = qg.type == PIMC_GRID_LOG ? log(qg.spacing) : 0.0;
  F.x = (const double*)(base + off_q);
  F.nt = nt;
  auto grid_dev = [](const HostGrid& g, const unsigned char* x, const unsigned char* iv) {
    const bool lg = g.type == PIMC_GRID_LOG;
    return GridDev{g.type, g.n, g.start, g.end, g.deltainv, g.r0, lg ? log(g.spacing) : 0.0, (const double*)x,
                   (const d4v*)iv, lg ? 1.0 / g.r0 : 0.0, lg ? (float)(1.0 / log2(g.spacing)) : 0.0f};
  };
  F.tgrid = grid_dev(tg, base + off_t, base + off_ti);
  F.qgrid = grid_dev(qg, base + off_q, base + off_qi);
  F.bc = (const d4v*)(base + off_b);
  F.smallest_beta = d->smallest_beta;
  F.z1z2 = d->z1z2;
  f->num_terms = -1;
  ctx->fits.push_back(std::move(f));
  ctx->fits_dirty = true;
  *fit_id = (int32_t)ctx->fits.size() - 1;
  return PIMC_OK;
}

int pimc_replica_set_ulong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, int32_t num_levels,
                           double grid_start, double grid_end, int32_t num_points, const double* coefs,
                           const double* dcoefs) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || num_levels < 1 || num_levels > PIMC_MAX_LEVELS || num_points < 2 || !coefs || !(grid_start < grid_end))
    return fail(PIMC_ERR_ARG, "replica_set_ulong: bad argument");
  RepCore& R = H->dev.core;
  if (sa < 0 || sb < 0 || sa >= R.ns || sb >= R.ns) return fail(PIMC_ERR_ARG, "replica_set_ulong: bad species");
  const double delta = (grid_end - grid_start) / (double)(num_points - 1);
  if ((R.ulong_n != 0 && (R.ulong_n != num_points || R.ulong_start != grid_start || R.ulong_delta != delta)) ||
      (R.ulong && R.ulong_levels != num_levels))
    return fail(PIMC_ERR_ARG, "replica_set_ulong: every pair of a replica needs the same grid and levels");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const size_t per_pair = (size_t)num_levels * num_points * 6;
  if (!R.ulong) {
    int rc = H->ulong.ensure((size_t)R.ns * R.ns * per_pair * sizeof(double));
    if (rc) return rc;
    HIP_TRY(hipMemset(H->ulong.p, 0, (size_t)R.ns * R.ns * per_pair * sizeof(double)));
    R.ulong_n = num_points;
    R.ulong_levels = num_levels;
    R.ulong_start = grid_start;
    R.ulong_delta = delta;
    R.ulong_deltainv = 1.0 / delta;
    R.ulong = H->ulong.as<double>();
  }
  if (!R.ulong32) {
    const size_t n32 = (size_t)R.ns * R.ns * num_levels * num_points;
    int rc = H->ulong32.ensure(n32 * sizeof(UlongKnot32));
    if (rc) return rc;
    HIP_TRY(hipMemset(H->ulong32.p, 0, n32 * sizeof(UlongKnot32)));
    R.ulong32 = H->ulong32.as<UlongKnot32>();
  }
  // the mixed path's knots: Y in FP64, B..F in FP32
  std::vector<UlongKnot32> c32((size_t)num_levels * num_points);
  for (size_t k = 0; k < c32.size(); k++) {
    c32[k].y = coefs[6 * k];
    for (int c = 0; c < 5; c++) c32[k].c[c] = (float)coefs[6 * k + 1 + c];
    c32[k].pad = 0.0f;
  }
  for (int pr : {sa * R.ns + sb, sb * R.ns + sa}) {
    HIP_TRY(hipMemcpy(H->ulong.as<double>() + pr * per_pair, coefs, per_pair * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(H->ulong32.as<UlongKnot32>() + (size_t)pr * num_levels * num_points, c32.data(),
                      c32.size() * sizeof(UlongKnot32), hipMemcpyHostToDevice));
    R.ulong_mask |= 1u << pr;
  }
  if (dcoefs) {
    if (!R.dulong) {
      int rc = H->dulong.ensure((size_t)R.ns * R.ns * per_pair * sizeof(double));
      if (rc) return rc;
      HIP_TRY(hipMemset(H->dulong.p, 0, (size_t)R.ns * R.ns * per_pair * sizeof(double)));
      R.dulong = H->dulong.as<double>();
    }
    for (int pr : {sa * R.ns + sb, sb * R.ns + sa}) {
      HIP_TRY(hipMemcpy(H->dulong.as<double>() + pr * per_pair, dcoefs, per_pair * sizeof(double),
                        hipMemcpyHostToDevice));
      H->dulong_mask |= 1u << pr;
    }
  }
  ctx->reps_dirty = true;
  return PIMC_OK;
}

int pimc_replica_set_vlong(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, const double* vlong_k,
                           double vlong_r0, double vshort_k0, double vlong_k0, double grid_start, double grid_end,
                           int32_t num_points, const double* coefs) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !coefs || num_points < 2 || !(grid_start < grid_end)) return fail(PIMC_ERR_ARG, "replica_set_vlong: bad argument");
  RepCore& R = H->dev.core;
  if (sa < 0 || sb < 0 || sa >= R.ns || sb >= R.ns) return fail(PIMC_ERR_ARG, "replica_set_vlong: bad species");
  if (R.nk > 0 && !vlong_k) return fail(PIMC_ERR_ARG, "replica_set_vlong: vlong_k is NULL");
  const double delta = (grid_end - grid_start) / (double)(num_points - 1);
  if (R.ulong_n != 0 && (R.ulong_n != num_points || R.ulong_start != grid_start || R.ulong_delta != delta))
    return fail(PIMC_ERR_ARG, "replica_set_vlong: the grid must match the replica's Ulong grid");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const int npair = R.ns * R.ns;
  const size_t per_pair = (size_t)num_points * 6;
  int rc;
  if (!R.vlong) {
    if ((rc = H->vlong.ensure((size_t)npair * per_pair * sizeof(double))) ||
        (rc = H->vlong_k.ensure((size_t)npair * std::max(1, R.nk) * sizeof(double))) ||
        (rc = H->vlr_const.ensure((size_t)npair * 3 * sizeof(double))))
      return rc;
    HIP_TRY(hipMemset(H->vlong.p, 0, (size_t)npair * per_pair * sizeof(double)));
    HIP_TRY(hipMemset(H->vlong_k.p, 0, (size_t)npair * std::max(1, R.nk) * sizeof(double)));
    HIP_TRY(hipMemset(H->vlr_const.p, 0, (size_t)npair * 3 * sizeof(double)));
    R.vlong = H->vlong.as<double>();
    R.vlong_k = H->vlong_k.as<double>();
    R.vlr_const = H->vlr_const.as<double>();
  }
  if (R.ulong_n == 0) {
    R.ulong_n = num_points;
    R.ulong_start = grid_start;
    R.ulong_delta = delta;
    R.ulong_deltainv = 1.0 / delta;
  }
  const double cst[3] = {vlong_r0, vshort_k0, vlong_k0};
  for (int pr : {sa * R.ns + sb, sb * R.ns + sa}) {
    HIP_TRY(hipMemcpy(H->vlong.as<double>() + pr * per_pair, coefs, per_pair * sizeof(double), hipMemcpyHostToDevice));
    if (R.nk > 0)
      HIP_TRY(hipMemcpy(H->vlong_k.as<double>() + (size_t)pr * R.nk, vlong_k, (size_t)R.nk * sizeof(double),
                        hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(H->vlr_const.as<double>() + (size_t)pr * 3, cst, sizeof cst, hipMemcpyHostToDevice));
    R.vlong_mask |= 1u << pr;
  }
  ctx->reps_dirty = true;
  return PIMC_OK;
}

int pimc_fit_eval_u(pimc_ctx* ctx, int32_t fit_id, int32_t level, int64_t n, const double* q, const double* z,
                    const double* s2, int32_t precision, double* out) {
  if (!ctx || fit_id < 0 || fit_id >= (int)ctx->fits.size()) return fail(PIMC_ERR_ARG, "fit_eval_u: bad fit");
  if (n < 0 || (n > 0 && (!q || !z || !s2 || !out))) return fail(PIMC_ERR_ARG, "fit_eval_u: bad arrays");
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "fit_eval_u: bad precision");
  const FitDev& F = ctx->fits[fit_id]->dev;
  if (level < 0 || level + F.taupos >= F.num_tau) return fail(PIMC_ERR_ARG, "fit_eval_u: level out of range");
  if (n == 0) return PIMC_OK;
  const int nt_launch = std::max(0, ctx->fits[fit_id]->num_terms);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  size_t b = (size_t)n * sizeof(double);
  if ((rc = ctx->d_in0.ensure(b)) || (rc = ctx->d_in1.ensure(b)) || (rc = ctx->d_in2.ensure(b)) ||
      (rc = ctx->d_out.ensure(b)))
    return rc;
  HIP_TRY(hipMemcpyAsync(ctx->d_in0.p, q, b, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->d_in1.p, z, b, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ctx->d_in2.p, s2, b, hipMemcpyHostToDevice, ctx->stream));
  LaunchFitEval a{ctx->d_fits.as<FitDev>(), fit_id, level, n, ctx->d_in0.as<double>(), ctx->d_in1.as<double>(),
                  ctx->d_in2.as<double>(), ctx->d_out.as<double>()};
  launch_fit_eval(nt_launch, precision == PIMC_MIXED, a, ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->d_out.p, b, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

int pimc_replica_create(pimc_ctx* ctx, const pimc_replica_desc* d, int32_t* replica_id) {
  if (!ctx || !d || !replica_id) return fail(PIMC_ERR_ARG, "replica_create: NULL argument");
  if (d->num_particles < 1 || d->num_time_slices < 1 || d->num_species < 1 || d->num_species > PIMC_MAX_SPECIES)
    return fail(PIMC_ERR_ARG, "replica_create: bad sizes");
  if (!d->species_count || !d->lambda || !d->pair_fit) return fail(PIMC_ERR_ARG, "replica_create: NULL arrays");
  if (d->num_images < 0) return fail(PIMC_ERR_ARG, "replica_create: bad num_images");
  auto H = std::make_unique<HostReplica>();
  RepCore& R = H->dev.core;
  RepSpecies& S = H->dev.sp;
  R.N = d->num_particles;
  R.nslices = d->num_time_slices + 1;
  R.ns = d->num_species;
  R.num_images = d->num_images;
  R.tau = d->tau;
  for (int i = 0; i < 3; i++) {
    if (!(d->box[i] > 0)) return fail(PIMC_ERR_ARG, "replica_create: box must be positive");
    R.box[i] = d->box[i];
    R.boxinv[i] = 1.0 / d->box[i];
    R.isper[i] = d->periodic[i] ? 1.0 : 0.0;
  }
  int tot = 0;
  for (int s = 0; s < R.ns; s++) {
    if (d->species_count[s] < 0) return fail(PIMC_ERR_ARG, "replica_create: negative species count");
    S.first[s] = tot;
    S.count[s] = d->species_count[s];
    S.lambda[s] = d->lambda[s];
    tot += d->species_count[s];
  }
  if (tot != R.N) return fail(PIMC_ERR_ARG, "replica_create: species counts do not sum to num_particles");
  for (int i = 0; i < R.ns * R.ns; i++) {
    if (d->pair_fit[i] < 0 || d->pair_fit[i] >= (int)ctx->fits.size())
      return fail(PIMC_ERR_ARG, "replica_create: bad pair fit id");
    S.fit[i] = d->pair_fit[i];
  }
  R.stream = d->rng_stream;
  H->M = d->num_time_slices;
  R.nk = 0;
  H->use_long_range = d->use_long_range != 0;
  R.use_long_range = H->use_long_range ? 1 : 0;
  R.use_background = d->use_background != 0;
  R.lr_on = 0;
  R.lr_mask = 0;
  R.lr_levels = 0;
  if (d->k_cutoff > 0) {
    setup_kvecs(*H, d->k_cutoff);
    if (R.maxk0 > 64 || R.maxk1 > 64 || R.maxk2 > 64)
      return fail(PIMC_ERR_UNSUPPORTED, "replica_create: k_cutoff needs more than 64 k indices per dimension");
  }
  HIP_TRY(hipSetDevice(ctx->device));
  size_t nb = (size_t)R.nslices * 3 * R.N * sizeof(double);
  int rc;
  if ((rc = H->path0.ensure(nb)) || (rc = H->path1.ensure(nb)) || (rc = H->perm0.ensure(R.N * sizeof(int))) ||
      (rc = H->perm1.ensure(R.N * sizeof(int))) || (rc = H->joinbuf.ensure(sizeof(int))))
    return rc;
  HIP_TRY(hipMemset(H->joinbuf.p, 0, sizeof(int)));
  HIP_TRY(hipMemset(H->path0.p, 0, nb));
  HIP_TRY(hipMemset(H->path1.p, 0, nb));
  std::vector<int> id(R.N);
  for (int p = 0; p < R.N; p++) id[p] = p;
  HIP_TRY(hipMemcpy(H->perm0.p, id.data(), R.N * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(H->perm1.p, id.data(), R.N * sizeof(int), hipMemcpyHostToDevice));
  R.path0 = H->path0.as<double>();
  R.path1 = H->path1.as<double>();
  R.perm0 = H->perm0.as<int>();
  R.perm1 = H->perm1.as<int>();
  R.join = H->joinbuf.as<int>();
  if (R.nk > 0) {
    size_t nr = (size_t)R.nslices * R.ns * R.nk * sizeof(d2v);
    if ((rc = H->rho0.ensure(nr)) || (rc = H->rho1.ensure(nr)) || (rc = H->kidx_dev.ensure((size_t)R.nk * 4 * sizeof(int))))
      return rc;
    HIP_TRY(hipMemset(H->rho0.p, 0, nr));
    HIP_TRY(hipMemset(H->rho1.p, 0, nr));
    std::vector<int> k4((size_t)R.nk * 4, 0);
    for (int k = 0; k < R.nk; k++)
      for (int d = 0; d < 3; d++) k4[4 * k + d] = H->kidx[3 * k + d];
    HIP_TRY(hipMemcpy(H->kidx_dev.p, k4.data(), k4.size() * sizeof(int), hipMemcpyHostToDevice));
    R.rho0 = H->rho0.as<d2v>();
    R.rho1 = H->rho1.as<d2v>();
    R.kidx = H->kidx_dev.as<int>();
  }
  ctx->reps.push_back(std::move(H));
  ctx->reps_dirty = true;
  *replica_id = (int32_t)ctx->reps.size() - 1;
  return PIMC_OK;
}

int pimc_replica_num_kvecs(pimc_ctx* ctx, int32_t replica, int32_t* nk) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !nk) return fail(PIMC_ERR_ARG, "replica_num_kvecs: bad argument");
  *nk = H->dev.core.nk;
  return PIMC_OK;
}
int pimc_replica_kvecs(pimc_ctx* ctx, int32_t replica, double* kvecs, int32_t* kidx) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H) return fail(PIMC_ERR_ARG, "replica_kvecs: bad replica");
  if (kvecs && !H->kvecs.empty()) memcpy(kvecs, H->kvecs.data(), H->kvecs.size() * sizeof(double));
  if (kidx && !H->kidx.empty()) memcpy(kidx, H->kidx.data(), H->kidx.size() * sizeof(int32_t));
  return PIMC_OK;
}
int pimc_replica_set_long_range(pimc_ctx* ctx, int32_t replica, int32_t sa, int32_t sb, int32_t num_levels,
                                const double* ulong_k, const double* ulong_r0, const double* ushort_k0,
                                const double* ulong_k0, const double* dulong_k, const double* dulong_r0,
                                const double* dushort_k0, const double* dulong_k0) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || num_levels < 1 || num_levels > PIMC_MAX_LEVELS || !ulong_k || !ulong_r0 || !ushort_k0 || !ulong_k0)
    return fail(PIMC_ERR_ARG, "replica_set_long_range: bad argument");
  const int nd = !!dulong_k + !!dulong_r0 + !!dushort_k0 + !!dulong_k0;
  if (nd != 0 && nd != 4) return fail(PIMC_ERR_ARG, "replica_set_long_range: give all four derivative tables or none");
  RepCore& R = H->dev.core;
  if (sa < 0 || sb < 0 || sa >= R.ns || sb >= R.ns) return fail(PIMC_ERR_ARG, "replica_set_long_range: bad species");
  if (R.nk == 0) return fail(PIMC_ERR_STATE, "replica_set_long_range: replica has no k set");
  if (R.lr_levels != 0 && R.lr_levels != num_levels)
    return fail(PIMC_ERR_ARG, "replica_set_long_range: every pair needs the same number of levels");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const int npair = R.ns * R.ns;
  if (R.lr_levels == 0) {
    int rc;
    size_t bk = (size_t)npair * num_levels * R.nk * sizeof(double), bc = (size_t)npair * 3 * num_levels * sizeof(double);
    if ((rc = H->ulong_k.ensure(bk)) || (rc = H->lr_const.ensure(bc))) return rc;
    HIP_TRY(hipMemset(H->ulong_k.p, 0, bk));
    HIP_TRY(hipMemset(H->lr_const.p, 0, bc));
    R.lr_levels = num_levels;
    R.ulong_k = H->ulong_k.as<double>();
    R.lr_const = H->lr_const.as<double>();
  }
  for (int pr : {sa * R.ns + sb, sb * R.ns + sa}) {
    HIP_TRY(hipMemcpy(H->ulong_k.as<double>() + (size_t)pr * num_levels * R.nk, ulong_k,
                      (size_t)num_levels * R.nk * sizeof(double), hipMemcpyHostToDevice));
    double* c = H->lr_const.as<double>() + (size_t)pr * 3 * num_levels;
    HIP_TRY(hipMemcpy(c, ulong_r0, num_levels * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c + num_levels, ushort_k0, num_levels * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c + 2 * num_levels, ulong_k0, num_levels * sizeof(double), hipMemcpyHostToDevice));
    R.lr_mask |= 1u << pr;
  }
  if (nd == 4) {
    const size_t bk = (size_t)npair * num_levels * R.nk * sizeof(double), bc = (size_t)npair * 3 * num_levels * sizeof(double);
    if (!R.dulong_k) {
      int rc;
      if ((rc = H->dulong_k.ensure(bk)) || (rc = H->dlr_const.ensure(bc))) return rc;
      HIP_TRY(hipMemset(H->dulong_k.p, 0, bk));
      HIP_TRY(hipMemset(H->dlr_const.p, 0, bc));
      R.dulong_k = H->dulong_k.as<double>();
      R.dlr_const = H->dlr_const.as<double>();
    }
    for (int pr : {sa * R.ns + sb, sb * R.ns + sa}) {
      HIP_TRY(hipMemcpy(H->dulong_k.as<double>() + (size_t)pr * num_levels * R.nk, dulong_k,
                        (size_t)num_levels * R.nk * sizeof(double), hipMemcpyHostToDevice));
      double* c = H->dlr_const.as<double>() + (size_t)pr * 3 * num_levels;
      HIP_TRY(hipMemcpy(c, dulong_r0, num_levels * sizeof(double), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(c + num_levels, dushort_k0, num_levels * sizeof(double), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(c + 2 * num_levels, dulong_k0, num_levels * sizeof(double), hipMemcpyHostToDevice));
      H->dlr_mask |= 1u << pr;
    }
  }
  R.lr_on = (H->use_long_range && R.lr_mask) ? 1 : 0;
  R.lr_terms = 0;
  for (int a = 0; a < R.ns; a++)
    for (int b = a; b < R.ns; b++) R.lr_terms += (R.lr_mask >> (a * R.ns + b)) & 1u;
  ctx->reps_dirty = true;
  return PIMC_OK;
}

int pimc_path_set(pimc_ctx* ctx, int32_t replica, int32_t mode, const double* beads) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !beads || mode < 0 || mode > 2) return fail(PIMC_ERR_ARG, "path_set: bad argument");
  const RepCore& R = H->dev.core;
  std::vector<double> tmp((size_t)R.nslices * 3 * R.N);
  to_device_layout(beads, tmp.data(), R.nslices, R.N);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  size_t nb = tmp.size() * sizeof(double);
  if (mode != 1) HIP_TRY(hipMemcpy(H->path0.p, tmp.data(), nb, hipMemcpyHostToDevice));
  if (mode != 0) HIP_TRY(hipMemcpy(H->path1.p, tmp.data(), nb, hipMemcpyHostToDevice));
  return PIMC_OK;
}
int pimc_path_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* beads) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !beads || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "path_get: bad argument");
  const RepCore& R = H->dev.core;
  std::vector<double> tmp((size_t)R.nslices * 3 * R.N);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(tmp.data(), mode ? H->path1.p : H->path0.p, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  to_host_layout(tmp.data(), beads, R.nslices, R.N);
  return PIMC_OK;
}
int pimc_path_set_beads(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t ptcl, int32_t slice1, int32_t slice2,
                        const double* xyz) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !xyz || mode < 0 || mode > 2 || ptcl < 0 || ptcl >= H->dev.core.N || slice1 < 0 || slice2 > H->M ||
      slice1 > slice2)
    return fail(PIMC_ERR_ARG, "path_set_beads: bad argument");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  const size_t bytes = (size_t)(slice2 - slice1 + 1) * 3 * sizeof(double);
  if ((rc = ctx->d_in2.ensure(bytes))) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->d_in2.p, xyz, bytes, hipMemcpyHostToDevice, ctx->stream));
  launch_set_beads(ctx->d_reps.as<ReplicaDev>(), replica, mode, ptcl, slice1, slice2, ctx->d_in2.as<double>(),
                   ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}
int pimc_permutation_get(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t* perm) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !perm || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "permutation_get: bad argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(perm, mode ? H->perm1.p : H->perm0.p, H->dev.core.N * sizeof(int), hipMemcpyDeviceToHost));
  return PIMC_OK;
}

static int accept_reject(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2, const int32_t* ptcls,
                         int32_t n, bool accept) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || (n > 0 && !ptcls) || n < 0 || n > PIMC_MAX_ACTIVE || slice1 < 0 || slice2 > H->M || slice1 > slice2)
    return fail(PIMC_ERR_ARG, "accept/reject copy: bad argument");
  for (int i = 0; i < n; i++)
    if (ptcls[i] < 0 || ptcls[i] >= H->dev.core.N) return fail(PIMC_ERR_ARG, "accept/reject copy: bad particle");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  launch_accept_reject(ctx->d_reps.as<ReplicaDev>(), replica, slice1, slice2, ptcls, n, accept, ctx->stream);
  HIP_TRY(hipGetLastError());
  const RepCore& R = H->dev.core;
  if (R.nk > 0) {
    if ((rc = refresh_rho1(ctx, H))) return rc;
    // Rho_k[OLD] <- Rho_k[NEW] (accept) or the reverse over the slice range, all species
    const size_t off = (size_t)slice1 * R.ns * R.nk, cnt = (size_t)(slice2 - slice1 + 1) * R.ns * R.nk;
    d2v* dst = accept ? H->rho0.as<d2v>() : H->rho1.as<d2v>();
    const d2v* src = accept ? H->rho1.as<d2v>() : H->rho0.as<d2v>();
    HIP_TRY(hipMemcpyAsync(dst + off, src + off, cnt * sizeof(d2v), hipMemcpyDeviceToDevice, ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}
int pimc_accept_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2, const int32_t* ptcls,
                     int32_t nptcls) {
  return accept_reject(ctx, replica, slice1, slice2, ptcls, nptcls, true);
}
int pimc_reject_copy(pimc_ctx* ctx, int32_t replica, int32_t slice1, int32_t slice2, const int32_t* ptcls,
                     int32_t nptcls) {
  return accept_reject(ctx, replica, slice1, slice2, ptcls, nptcls, false);
}
int pimc_move_join(pimc_ctx* ctx, int32_t replica, int32_t old_join, int32_t new_join) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || old_join < 0 || new_join < 0 || old_join > H->M || new_join > H->M)
    return fail(PIMC_ERR_ARG, "move_join: bad argument");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  launch_move_join(ctx->d_reps.as<ReplicaDev>(), replica, H->dev.core.N, old_join, new_join, ctx->stream);
  HIP_TRY(hipGetLastError());
  const int32_t join = new_join;   // PathDataClass::MoveJoin: Join = newJoin
  HIP_TRY(hipMemcpyAsync(H->joinbuf.p, &join, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}
int pimc_replica_join(pimc_ctx* ctx, int32_t replica, int32_t* join) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !join) return fail(PIMC_ERR_ARG, "replica_join: bad argument");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(join, H->joinbuf.p, sizeof(int32_t), hipMemcpyDeviceToHost));
  return PIMC_OK;
}
int pimc_put_in_box(pimc_ctx* ctx, int32_t replica, int32_t fast, int64_t n, double* v) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || n < 0 || (n > 0 && !v) || (fast != 0 && fast != 1)) return fail(PIMC_ERR_ARG, "put_in_box: bad argument");
  if (n == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  const size_t bytes = (size_t)n * 3 * sizeof(double);
  if ((rc = ctx->d_in0.ensure(bytes))) return rc;
  HIP_TRY(hipMemcpyAsync(ctx->d_in0.p, v, bytes, hipMemcpyHostToDevice, ctx->stream));
  launch_put_in_box(ctx->d_reps.as<ReplicaDev>(), replica, fast, n, ctx->d_in0.as<double>(), ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(v, ctx->d_in0.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

int pimc_permute_htable(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t species, int32_t slice1,
                        int32_t slice2, double epsilon, float* htable) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !htable || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "permute_htable: bad argument");
  const RepCore& R = H->dev.core;
  if (species < 0 || species >= R.ns) return fail(PIMC_ERR_ARG, "permute_htable: bad species");
  if (slice1 < 0 || slice2 <= slice1 || slice2 >= R.nslices) return fail(PIMC_ERR_ARG, "permute_htable: bad slices");
  if (!(epsilon > 0.0)) return fail(PIMC_ERR_ARG, "permute_htable: epsilon must be positive");
  const int first = H->dev.sp.first[species], n = H->dev.sp.count[species];
  // PermuteTableClass.cc:124-129
  const double beta = R.tau * (double)(slice2 - slice1);
  const double four_lambda_beta_inv = 1.0 / (4.0 * H->dev.sp.lambda[species] * beta);
  const double log_eps = log(epsilon);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  const size_t bytes = (size_t)n * n * sizeof(float);
  DevBuf out;
  if ((rc = out.ensure(bytes))) return rc;
  launch_permute_htable(ctx->d_reps.as<ReplicaDev>(), replica, mode, first, n, slice1, slice2, four_lambda_beta_inv,
                        log_eps, out.as<float>(), ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(htable, out.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

int pimc_cycle_apply(pimc_ctx* ctx, int32_t replica, int32_t first_ptcl, int32_t slice, int32_t length,
                     const int32_t* cycle) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !cycle || length < 1 || length > 4) return fail(PIMC_ERR_ARG, "cycle_apply: bad argument");
  const RepCore& R = H->dev.core;
  if (slice < 0 || slice >= R.nslices || first_ptcl < 0) return fail(PIMC_ERR_ARG, "cycle_apply: bad slice or first_ptcl");
  int ptcls[4];
  for (int i = 0; i < length; i++) {
    ptcls[i] = first_ptcl + cycle[i];
    if (cycle[i] < 0 || ptcls[i] >= R.N) return fail(PIMC_ERR_ARG, "cycle_apply: particle out of range");
    for (int k = 0; k < i; k++)
      if (ptcls[k] == ptcls[i]) return fail(PIMC_ERR_ARG, "cycle_apply: repeated particle");
  }
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  launch_cycle_apply(ctx->d_reps.as<ReplicaDev>(), replica, slice, ptcls, length, ctx->stream);
  HIP_TRY(hipGetLastError());
  if (length > 1) H->maybe_permuted = true;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

int pimc_replica_set_permute(pimc_ctx* ctx, int32_t replica, int32_t species, const double* gamma,
                             double epsilon) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H) return fail(PIMC_ERR_ARG, "replica_set_permute: bad replica");
  if (!gamma) {
    H->permute_on = false;
    return PIMC_OK;
  }
  if (species < 0 || species >= H->dev.core.ns) return fail(PIMC_ERR_ARG, "replica_set_permute: bad species");
  if (!(epsilon > 0.0)) return fail(PIMC_ERR_ARG, "replica_set_permute: epsilon must be positive");
  for (int i = 0; i < 4; i++)
    if (!(gamma[i] > 0.0)) return fail(PIMC_ERR_ARG, "replica_set_permute: gamma must be positive");
  H->permute_on = true;
  H->permute_species = species;
  for (int i = 0; i < 4; i++) H->permute.gamma[i] = gamma[i];
  H->permute.epsilon = epsilon;
  return PIMC_OK;
}

int pimc_rho_k_update_all(pimc_ctx* ctx, int32_t replica) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H) return fail(PIMC_ERR_ARG, "rho_k_update_all: bad replica");
  const RepCore& R = H->dev.core;
  if (R.nk == 0) return fail(PIMC_ERR_STATE, "rho_k_update_all: replica has no k set");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  launch_rho_all(ctx->d_reps.as<ReplicaDev>(), replica, R.nslices, R.ns, R.nk,
                 2 * (R.maxk0 + R.maxk1 + R.maxk2) + 3, ctx->stream);
  HIP_TRY(hipGetLastError());
  H->rho1_stale = false;   // both copies written
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}
int pimc_rho_k_get(pimc_ctx* ctx, int32_t replica, int32_t mode, double* rho) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !rho || mode < 0 || mode > 1) return fail(PIMC_ERR_ARG, "rho_k_get: bad argument");
  const RepCore& R = H->dev.core;
  if (R.nk == 0) return fail(PIMC_ERR_STATE, "rho_k_get: replica has no k set");
  HIP_TRY(hipSetDevice(ctx->device));
  if (mode == 1) {
    int rc = refresh_rho1(ctx, H);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipMemcpy(rho, mode ? H->rho1.p : H->rho0.p, (size_t)R.nslices * R.ns * R.nk * sizeof(d2v),
                    hipMemcpyDeviceToHost));
  return PIMC_OK;
}

int pimc_action_short_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                            double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_short_range: bad argument");
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "action_short_range: bad precision");
  if (nreq == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  CallIO io;
  rc = stage_call(ctx, reqs, nreq, io);
  if (rc) return rc;
  int maxN = 0, nt = -1;
  for (int i = 0; i < nreq; i++) {
    const ReplicaDev& R = ctx->reps[reqs[i].replica]->dev;
    maxN = std::max(maxN, R.core.N);
    for (int f = 0; f < R.core.ns * R.core.ns; f++) {
      const HostFit& F = *ctx->fits[R.sp.fit[f]];
      if (F.num_terms >= 0) {
        if (nt < 0) nt = F.num_terms;
        if (F.num_terms != nt)
          return fail(PIMC_ERR_UNSUPPORTED, "action_short_range: fits with different num_terms in one call");
      }
      if (reqs[i].level >= F.num_levels())
        return fail(PIMC_ERR_ARG, "action_short_range: level beyond the fit's tau levels");
      if (F.dev.kind == 1 && R.core.use_long_range &&
          (!((R.core.ulong_mask >> f) & 1u) || reqs[i].level >= R.core.ulong_levels))
        return fail(PIMC_ERR_STATE, "action_short_range: long-range pair fit without Ulong tables");
    }
  }
  if (nt < 0) nt = 0;
  const int ntiles = (maxN + kBlock - 1) / kBlock;
  const int nsplit = short_range_split(nreq, ntiles);
  if ((rc = ctx->d_partials.ensure((size_t)nreq * ntiles * nsplit * sizeof(double)))) return rc;
  LaunchShortRange a{ctx->d_reps.as<ReplicaDev>(), ctx->d_fits.as<FitDev>(), io.reqs, nreq, ntiles,
                     ctx->d_partials.as<double>(), io.out, nsplit};
  launch_short_range(nt, precision == PIMC_MIXED, a, ctx->stream);
  HIP_TRY(hipGetLastError());
  return finish_call(ctx, io, nreq, out);
}

int pimc_action_short_range_noalloc(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                                    double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out)))
    return fail(PIMC_ERR_ARG, "action_short_range_noalloc: bad argument");
  if (precision != PIMC_FP64 && precision != PIMC_MIXED)
    return fail(PIMC_ERR_ARG, "action_short_range_noalloc: bad precision");
  if (nreq > PIMC_NOALLOC_MAX_REQS)
    return fail(PIMC_ERR_ARG, "action_short_range_noalloc: more than PIMC_NOALLOC_MAX_REQS requests");
  if (nreq == 0) return PIMC_OK;
  unsigned char* base =
      ctx->device < kNoallocDevices ? (unsigned char*)g_noalloc_ptr[ctx->device] : nullptr;
  if (!base) return fail(PIMC_ERR_STATE, "action_short_range_noalloc: no scratch array on this device");
  if ((ctx->fits_dirty && !ctx->fits.empty()) || (ctx->reps_dirty && !ctx->reps.empty()))
    return fail(PIMC_ERR_STATE, "action_short_range_noalloc: fits or replicas changed since the last upload; "
                                "call another action entry point once first");
  ReqDev h[PIMC_NOALLOC_MAX_REQS];
  int maxN = 0, nt = -1;
  for (int i = 0; i < nreq; i++) {
    int rc = check_req(ctx, reqs[i]);
    if (rc) return rc;
    const ReplicaDev& R = ctx->reps[reqs[i].replica]->dev;
    maxN = std::max(maxN, R.core.N);
    for (int f = 0; f < R.core.ns * R.core.ns; f++) {
      const HostFit& F = *ctx->fits[R.sp.fit[f]];
      if (F.num_terms >= 0) {
        if (nt < 0) nt = F.num_terms;
        if (F.num_terms != nt)
          return fail(PIMC_ERR_UNSUPPORTED, "action_short_range_noalloc: fits with different num_terms in one call");
      }
      if (reqs[i].level >= F.num_levels())
        return fail(PIMC_ERR_ARG, "action_short_range_noalloc: level beyond the fit's tau levels");
      if (F.dev.kind == 1 && R.core.use_long_range &&
          (!((R.core.ulong_mask >> f) & 1u) || reqs[i].level >= R.core.ulong_levels))
        return fail(PIMC_ERR_STATE, "action_short_range_noalloc: long-range pair fit without Ulong tables");
    }
    h[i].replica = reqs[i].replica;
    h[i].mode = reqs[i].mode;
    h[i].slice1 = reqs[i].slice1;
    h[i].slice2 = reqs[i].slice2;
    h[i].level = reqs[i].level;
    h[i].nptcls = reqs[i].nptcls;
    for (int k = 0; k < PIMC_MAX_ACTIVE; k++) h[i].ptcls[k] = k < reqs[i].nptcls ? reqs[i].ptcls[k] : -1;
  }
  if (nt < 0) nt = 0;
  // scratch layout: requests [nreq], outputs [nreq], partial sums [nreq][ntiles * nsplit]
  const int ntiles = (maxN + kBlock - 1) / kBlock;
  const int nsplit = short_range_split(nreq, ntiles);
  const size_t req_bytes = ((size_t)nreq * sizeof(ReqDev) + 255) & ~(size_t)255;
  const size_t out_bytes = ((size_t)nreq * sizeof(double) + 255) & ~(size_t)255;
  const size_t part_bytes = (size_t)nreq * ntiles * nsplit * sizeof(double);
  if (req_bytes + out_bytes + part_bytes > PIMC_NOALLOC_SCRATCH_BYTES)
    return fail(PIMC_ERR_ARG, "action_short_range_noalloc: requests x particle tiles exceed the scratch array");
  ReqDev* d_reqs = (ReqDev*)base;
  double* d_out = (double*)(base + req_bytes);
  double* d_part = (double*)(base + req_bytes + out_bytes);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(d_reqs, h, (size_t)nreq * sizeof(ReqDev), hipMemcpyHostToDevice, ctx->stream));
  LaunchShortRange a{ctx->d_reps.as<ReplicaDev>(), ctx->d_fits.as<FitDev>(), d_reqs, nreq, ntiles, d_part, d_out,
                     nsplit};
  launch_short_range(nt, precision == PIMC_MIXED, a, ctx->stream);   // every tile writes its partial sum
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, d_out, (size_t)nreq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));   // h and the caller's buffers are read or written until here
  return PIMC_OK;
}

int pimc_action_kinetic(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_kinetic: bad argument");
  if (nreq == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  CallIO io;
  if ((rc = stage_call(ctx, reqs, nreq, io))) return rc;
  launch_kinetic(ctx->d_reps.as<ReplicaDev>(), io.reqs, nreq, io.out, ctx->stream);
  HIP_TRY(hipGetLastError());
  return finish_call(ctx, io, nreq, out);
}

int pimc_action_long_range(pimc_ctx* ctx, const pimc_action_req* reqs, int32_t nreq, int32_t precision,
                           double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_long_range: bad argument");
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "action_long_range: bad precision");
  if (nreq == 0) return PIMC_OK;
  std::vector<int> rids(nreq), nss(nreq), modes(nreq);
  int max_nk = 0;
  for (int i = 0; i < nreq; i++) {
    int rc = check_req(ctx, reqs[i]);
    if (rc) return rc;
    const RepCore& R = ctx->reps[reqs[i].replica]->dev.core;
    if (R.nk == 0) return fail(PIMC_ERR_STATE, "action_long_range: replica has no k set");
    if (R.lr_mask && reqs[i].level >= R.lr_levels)
      return fail(PIMC_ERR_ARG, "action_long_range: level beyond the long-range tables");
    if (reqs[i].mode == 1 && !long_range_phase_fits(2 * (R.maxk0 + R.maxk1 + R.maxk2) + 3, reqs[i].nptcls))
      return fail(PIMC_ERR_UNSUPPORTED, "action_long_range: the k set is too large for the moved particles' phase rows");
    rids[i] = reqs[i].replica;
    nss[i] = R.ns;
    modes[i] = reqs[i].mode;
    max_nk = std::max(max_nk, R.nk);
  }
  const int kblocks = long_range_kblocks(max_nk);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  for (int i = 0; i < nreq; i++)
    if ((rc = refresh_rho1(ctx, ctx->reps[reqs[i].replica].get()))) return rc;
  CallIO io;
  if ((rc = stage_call(ctx, reqs, nreq, io))) return rc;
  if ((rc = ctx->d_partials.ensure((size_t)nreq * kblocks * sizeof(double)))) return rc;
  const std::vector<LrGroup> groups = long_range_groups(nreq, rids.data(), nss.data(), modes.data());
  std::pair<hipEvent_t, hipEvent_t> ev;   // device time, read by pimc_bisection_stats_take
  if (int rc_ev = timed_pair(ctx, ev)) return rc_ev;
  HIP_TRY(hipEventRecord(ev.first, ctx->stream));
  launch_long_range(precision == PIMC_MIXED, ctx->d_reps.as<ReplicaDev>(), io.reqs, nreq, groups, kblocks,
                    ctx->d_partials.as<double>(), io.out, ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev.second, ctx->stream));
  return finish_call(ctx, io, nreq, out);
}

int pimc_dist_disp(pimc_ctx* ctx, int32_t replica, int32_t mode, int64_t n, const int32_t* idx, double* out,
                   int32_t* images) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !idx || !out || mode < 0 || mode > 1 || n < 0) return fail(PIMC_ERR_ARG, "dist_disp: bad argument");
  for (int64_t i = 0; i < n; i++) {
    const int32_t* q = idx + 4 * i;
    if (q[0] < 0 || q[0] > H->M || q[1] < 0 || q[1] > H->M || q[2] < 0 || q[2] >= H->dev.core.N || q[3] < 0 ||
        q[3] >= H->dev.core.N)
      return fail(PIMC_ERR_ARG, "dist_disp: index out of range");
  }
  if (n == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  if ((rc = ctx->d_idx.ensure((size_t)n * 4 * sizeof(int))) || (rc = ctx->d_out.ensure((size_t)n * 8 * sizeof(double))) ||
      (rc = ctx->d_img.ensure((size_t)n * 6 * sizeof(int))))
    return rc;
  HIP_TRY(hipMemcpyAsync(ctx->d_idx.p, idx, (size_t)n * 4 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  launch_dist_disp(ctx->d_reps.as<ReplicaDev>(), replica, mode, n, ctx->d_idx.as<int>(), ctx->d_out.as<double>(),
                   ctx->d_img.as<int>(), ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, ctx->d_out.p, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (images)
    HIP_TRY(hipMemcpyAsync(images, ctx->d_img.p, (size_t)n * 6 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PIMC_OK;
}

}  // extern "C"

namespace {

// LDS for the staged long-range phase tables of a bisection launch.
constexpr size_t kLrPhaseBytes = 24 * 1024;
// Cycle-table entries one block of a TABLE launch stores (8 B each); larger tables are walked
// by one thread instead.
constexpr int kCycleScratchEntries = 1 << 16;

// What a bisection launch over a replica list needs, validated.
struct BisectionPrep {
  int nt = 0;
  bool lean = false;
  bool table = false;        // PermuteType TABLE moves (k_bisection_table)
  size_t htable_bytes = 0;   // the largest moved species' HTable
  size_t table_bytes = 0, lr_bytes = 0;
  int lr_chunk = 1;
  std::vector<int> order;    // list positions, longest replica first (persistent launches)
  std::vector<int> ids;      // replica ids in launch order
};

int prepare_bisection(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                      int32_t precision, BisectionPrep& P) {
  if (precision != PIMC_FP64 && precision != PIMC_MIXED) return fail(PIMC_ERR_ARG, "bisection: bad precision");
  if (desc->num_levels < 1 || desc->num_levels > PIMC_MAX_LEVELS || desc->lowest_level < 0 ||
      desc->lowest_level >= desc->num_levels)
    return fail(PIMC_ERR_ARG, "bisection: bad levels");
  std::set<int> seen;
  int nt = -1;
  size_t table_bytes = 0, table_bytes_lin = 0, lr_bytes = 0;
  bool has_bc = false, has_lr = false;
  for (int r = 0; r < nrep; r++) {
    HostReplica* H = ctx->rep(replicas[r]);
    if (!H) return fail(PIMC_ERR_ARG, "bisection: bad replica");
    if (!seen.insert(replicas[r]).second) return fail(PIMC_ERR_ARG, "bisection: replica listed twice");
    const RepCore& R = H->dev.core;
    if ((1 << desc->num_levels) >= R.nslices) return fail(PIMC_ERR_ARG, "bisection: too few slices for num_levels");
    if (R.lr_on) {
      if (desc->lowest_level >= R.lr_levels) return fail(PIMC_ERR_ARG, "bisection: long-range tables lack the lowest level");
      H->rho1_stale = true;   // the chain updates rho_k OLD only
      // per slice: the NEW and OLD phase rows, FP64 (16 B) or FP32 (8 B) entries
      lr_bytes = std::max(lr_bytes, (size_t)2 * (2 * (R.maxk0 + R.maxk1 + R.maxk2) + 3) *
                                        (precision == PIMC_MIXED ? sizeof(f2v) : sizeof(d2v)));
      has_lr = true;
    }
    if (desc->species < 0 || desc->species >= R.ns || H->dev.sp.count[desc->species] < 1)
      return fail(PIMC_ERR_ARG, "bisection: bad species");
    const bool table = H->permute_on && H->permute_species == desc->species;
    if (r > 0 && table != P.table)
      return fail(PIMC_ERR_UNSUPPORTED, "bisection: replicas with and without TABLE permutations in one call");
    P.table = table;
    if (table) {
      if (R.lr_on) return fail(PIMC_ERR_UNSUPPORTED, "bisection: TABLE permutation moves with the long-range action");
      if (ctx->plan.cluster > 1) return fail(PIMC_ERR_UNSUPPORTED, "bisection: TABLE permutation moves need cluster = 1");
      if (desc->num_levels > 5) return fail(PIMC_ERR_UNSUPPORTED, "bisection: TABLE permutation moves need num_levels <= 5");
      const size_t n = (size_t)H->dev.sp.count[desc->species];
      P.htable_bytes = std::max(P.htable_bytes, n * n * sizeof(float));
    }
    if (ctx->plan.cluster > 1 && H->maybe_permuted) {
      // the blocks of a cluster cannot move the join of a permuted path together
      std::vector<int> p0(R.N), p1(R.N);
      HIP_TRY(hipSetDevice(ctx->device));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      HIP_TRY(hipMemcpy(p0.data(), H->perm0.p, R.N * sizeof(int), hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(p1.data(), H->perm1.p, R.N * sizeof(int), hipMemcpyDeviceToHost));
      for (int p = 0; p < R.N; p++)
        if (p0[p] != p || p1[p] != p)
          return fail(PIMC_ERR_UNSUPPORTED, "bisection: a cluster plan cannot run a replica whose permutation is "
                                            "not the identity (use a plan with cluster = 1)");
      H->maybe_permuted = false;
    }
    for (int f = 0; f < R.ns * R.ns; f++) {
      const HostFit& F = *ctx->fits[H->dev.sp.fit[f]];
      if (F.num_terms >= 0) {
        if (nt < 0) nt = F.num_terms;
        if (F.num_terms != nt) return fail(PIMC_ERR_UNSUPPORTED, "bisection: fits with different num_terms");
        table_bytes = std::max(table_bytes, precision == PIMC_MIXED ? F.staged32 : F.staged64);
        table_bytes_lin = std::max(table_bytes_lin, precision == PIMC_MIXED ? F.staged32_lin : F.staged64_lin);
      }
      if (desc->num_levels > F.num_levels()) return fail(PIMC_ERR_ARG, "bisection: pair fit has too few tau levels");
      has_bc = has_bc || F.dev.kind == 1 || F.dev.grid_type != PIMC_GRID_LINEAR;
      if (F.dev.kind == 1 && R.use_long_range && (!((R.ulong_mask >> f) & 1u) || desc->num_levels > R.ulong_levels))
        return fail(PIMC_ERR_STATE, "bisection: long-range pair fit without Ulong tables");
    }
  }
  P.nt = nt < 0 ? 0 : nt;
  P.lean = !has_bc && !has_lr;
  P.table_bytes = P.lean ? table_bytes_lin : table_bytes;
  // phase tables of up to 24 KB of slices at a time (all W/2^lowest + 1 slices when they
  // fit), within the 64 KB a block may use next to the staged spline table
  const int nslc = (1 << (desc->num_levels - desc->lowest_level)) + 1;
  const size_t room = P.table_bytes + 8192 < 64 * 1024 ? 64 * 1024 - P.table_bytes - 8192 : 0;
  P.lr_chunk = lr_bytes ? std::max(1, std::min(nslc, (int)(std::min(kLrPhaseBytes, room) / lr_bytes))) : 1;
  P.lr_bytes = lr_bytes * P.lr_chunk;
  // longest first: the short-range work of a move grows with N, the long-range work with nk
  P.order.resize(nrep);
  for (int r = 0; r < nrep; r++) P.order[r] = r;
  if (ctx->plan.persistent) {
    auto cost = [&](int r) {
      const RepCore& R = ctx->rep(replicas[r])->dev.core;
      return (double)R.N + (R.lr_on ? 0.05 * R.nk : 0.0);
    };
    std::stable_sort(P.order.begin(), P.order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  }
  P.ids.resize(nrep);
  for (int i = 0; i < nrep; i++) P.ids[i] = replicas[P.order[i]];
  return PIMC_OK;
}

// Grows the per-replica counters to cover every replica, keeping their values.
int ensure_acc(pimc_ctx* ctx) {
  if ((int)ctx->reps.size() <= ctx->acc_cap) return PIMC_OK;
  if (ctx->graphs_alive())
    return fail(PIMC_ERR_STATE, "replicas were added while a sweep graph exists; destroy the graphs first");
  const int cap = std::max((int)ctx->reps.size(), 2 * ctx->acc_cap);
  DevBuf grown;
  int rc = grown.ensure((size_t)cap * 5 * sizeof(int64_t));
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(grown.p, 0, (size_t)cap * 5 * sizeof(int64_t), ctx->stream));
  if (ctx->acc_cap)
    HIP_TRY(hipMemcpyAsync(grown.p, ctx->d_acc.p, (size_t)ctx->acc_cap * 5 * sizeof(int64_t), hipMemcpyDeviceToDevice,
                           ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->d_acc.swap(grown);
  ctx->acc_cap = cap;
  return PIMC_OK;
}

// Launch arguments over device buffers that stay valid while `ids`/`pos` do.
LaunchBisection make_launch(pimc_ctx* ctx, const BisectionPrep& P, const pimc_bisection_desc* desc, int nrep,
                            uint32_t first_move, int num_moves, const int* d_ids, const int* d_pos, int64_t* d_stats,
                            pimc_move_record* d_records, int* d_queue, const uint32_t* d_first, int* d_bar = nullptr,
                            double* d_part = nullptr) {
  LaunchBisection a{};
  a.reps = ctx->d_reps.as<ReplicaDev>();
  a.fits = ctx->d_fits.as<FitDev>();
  a.rep_ids = d_ids;
  a.rep_pos = d_pos;
  a.nrep = nrep;
  a.num_levels = desc->num_levels;
  a.lowest_level = desc->lowest_level;
  a.species = desc->species;
  a.seed = desc->seed;
  a.first_move = first_move;
  a.num_moves = num_moves;
  a.records = d_records;
  a.stats = d_stats;
  a.acc = ctx->d_acc.as<int64_t>();
  a.lds_table_bytes = P.table_bytes;
  a.lds_lr_bytes = P.lr_bytes;
  a.lr_chunk = P.lr_chunk;
  a.link_split = ctx->plan.link_split;
  a.queue = ctx->plan.persistent ? d_queue : nullptr;
  a.first_move_dev = d_first;
  a.cluster = std::max(1, ctx->plan.cluster);
  a.bar = d_bar;
  a.part = d_part;
  a.abort_flag = ctx->d_abort.as<int>();
  a.wait_ticks = ctx->wait_ticks;
  return a;
}

// Device buffers of a cluster launch over nrep replicas with the context's plan.
int ensure_cluster_buffers(pimc_ctx* ctx, int nrep, DevBuf& bar, DevBuf& part) {
  const int c = std::max(1, ctx->plan.cluster);
  if (c == 1) return PIMC_OK;
  int rc = bar.ensure((size_t)nrep * sizeof(int));
  if (!rc) rc = part.ensure((size_t)nrep * 2 * c * 4 * sizeof(double));
  return rc;
}

// A cluster launch needs all nrep * cluster blocks resident at once.
int check_cluster_fits(pimc_ctx* ctx, const BisectionPrep& P, const LaunchBisection& a, int precision, int nrep) {
  if (a.cluster <= 1) return PIMC_OK;
  const int per_cu = bisection_blocks_per_cu(P.nt, precision == PIMC_MIXED, P.lean, ctx->plan.block_threads,
                                             bisection_lds_bytes_for(a), 2);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
  if ((int64_t)nrep * a.cluster > (int64_t)per_cu * prop.multiProcessorCount)
    return fail(PIMC_ERR_ARG, "bisection: replicas x cluster (" + std::to_string((int64_t)nrep * a.cluster) +
                                  " blocks) exceed the resident blocks of the device (" +
                                  std::to_string(per_cu * prop.multiProcessorCount) + ")");
  return PIMC_OK;
}

// Reports (and clears) a cluster barrier wait that timed out in an earlier launch.
int take_abort(pimc_ctx* ctx) {
  int flag = 0;
  HIP_TRY(hipMemcpy(&flag, ctx->d_abort.p, sizeof(int), hipMemcpyDeviceToHost));
  if (!flag) return PIMC_OK;
  HIP_TRY(hipMemset(ctx->d_abort.p, 0, sizeof(int)));
  return fail(PIMC_ERR_DEVICE, "bisection: a cluster launch was abandoned (a barrier wait timed out, or a replica's "
                                "permutation is not the identity); the replicas' paths are undefined");
}

int persistent_grid(pimc_ctx* ctx, const BisectionPrep& P, const LaunchBisection& a, int precision, int nrep) {
  const int per_cu = bisection_blocks_per_cu(P.nt, precision == PIMC_MIXED, P.lean, ctx->plan.block_threads,
                                             bisection_lds_bytes_for(a), 1);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess || per_cu <= 0) return nrep;
  return std::max(1, std::min(nrep, per_cu * prop.multiProcessorCount));
}

// The TABLE part of a launch: each list entry's cycle weights and cut-off; one block per
// replica at a time, no work queue; LDS per block: the CU's LDS shared by the blocks the
// kernel's registers allow; grid: the blocks resident at once; cycle-table scratch: one region
// per block.
int table_launch_setup(pimc_ctx* ctx, const BisectionPrep& P, int nrep, int precision, LaunchBisection& a) {
  int rc;
  std::vector<PermuteDev> pd(nrep);
  for (int i = 0; i < nrep; i++) pd[i] = ctx->rep(P.ids[i])->permute;
  if ((rc = ctx->d_permute.ensure((size_t)nrep * sizeof(PermuteDev)))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));   // earlier launches may still read the old entries
  HIP_TRY(hipMemcpy(ctx->d_permute.p, pd.data(), (size_t)nrep * sizeof(PermuteDev), hipMemcpyHostToDevice));
  a.permute = ctx->d_permute.as<PermuteDev>();
  a.lds_htable_bytes = P.htable_bytes;
  a.queue = nullptr;
  int cus = 0, lds_cu = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  HIP_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx->device));
  a.permute_lds_limit = bisection_table_lds_limit((size_t)lds_cu);
  const int per_cu = bisection_table_blocks_per_cu(P.nt, precision == PIMC_MIXED, P.lean, bisection_table_lds_bytes(a));
  a.permute_grid = std::max(1, std::min(nrep, std::max(1, per_cu) * std::max(1, cus)));
  a.permute_cap = kCycleScratchEntries;
  if ((rc = ctx->d_cycles.ensure(bisection_table_scratch_bytes(a.permute_cap, a.permute_grid)))) return rc;
  a.permute_scratch = ctx->d_cycles.p;
  if (bisection_table_lds_bytes(a) > 64 * 1024)
    return fail(PIMC_ERR_UNSUPPORTED, "bisection: the moved species' HTable and the staged spline level exceed "
                                      "the 64 KB of LDS of a TABLE permutation block (too many particles)");
  return PIMC_OK;
}

int launch_timed(pimc_ctx* ctx, const BisectionPrep& P, const LaunchBisection& a, int precision, int nrep) {
  std::pair<hipEvent_t, hipEvent_t> ev;
  if (int rc_ev = timed_pair(ctx, ev)) return rc_ev;
  if (P.table) {   // one block of 256 threads per replica
    HIP_TRY(hipEventRecord(ev.first, ctx->stream));
    hipError_t e = launch_bisection_table(P.nt, precision == PIMC_MIXED, P.lean, a, ctx->stream);
    if (e != hipSuccess) return fail(PIMC_ERR_DEVICE, std::string("bisection launch: ") + hipGetErrorString(e));
    HIP_TRY(hipEventRecord(ev.second, ctx->stream));
    return PIMC_OK;
  }
  const int grid = ctx->plan.persistent ? persistent_grid(ctx, P, a, precision, nrep) : nrep;
  int rc = check_cluster_fits(ctx, P, a, precision, nrep);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ev.first, ctx->stream));
  hipError_t e = launch_bisection(P.nt, precision == PIMC_MIXED, P.lean, a, ctx->plan.block_threads, grid, ctx->stream);
  if (e != hipSuccess) return fail(PIMC_ERR_DEVICE, std::string("bisection launch: ") + hipGetErrorString(e));
  HIP_TRY(hipEventRecord(ev.second, ctx->stream));
  return PIMC_OK;
}

}  // namespace

extern "C" {

int pimc_bisection_run(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                       int64_t first_move, int32_t num_moves, int32_t precision, pimc_move_record* records,
                       pimc_bisection_stats* stats) {
  if (!ctx || !desc || nrep < 0 || (nrep > 0 && !replicas) || num_moves < 0 || first_move < 0 ||
      first_move + num_moves > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "bisection_run: bad argument");
  BisectionPrep P;
  int rc = prepare_bisection(ctx, replicas, nrep, desc, precision, P);
  if (rc) return rc;
  if (stats) memset(stats, 0, sizeof *stats);
  if (nrep == 0 || num_moves == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = ctx->sync_descriptors())) return rc;
  if ((rc = ensure_acc(ctx))) return rc;
  if ((rc = ctx->d_stats.ensure((size_t)nrep * 5 * sizeof(int64_t)))) return rc;
  if (records && (rc = ctx->d_records.ensure((size_t)nrep * num_moves * sizeof(pimc_move_record)))) return rc;
  if ((rc = ctx->d_queue.ensure(sizeof(int)))) return rc;
  if ((rc = ensure_cluster_buffers(ctx, nrep, ctx->d_bar, ctx->d_part))) return rc;
  if (ctx->h_repids != P.ids || ctx->h_reppos != P.order) {
    // the launch list is uploaded once and reused while it does not change
    if ((rc = ctx->d_repids.ensure((size_t)nrep * sizeof(int))) || (rc = ctx->d_reppos.ensure((size_t)nrep * sizeof(int))))
      return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));   // earlier launches may still read the old list
    ctx->h_repids = P.ids;
    ctx->h_reppos = P.order;
    HIP_TRY(hipMemcpy(ctx->d_repids.p, P.ids.data(), (size_t)nrep * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->d_reppos.p, P.order.data(), (size_t)nrep * sizeof(int), hipMemcpyHostToDevice));
  }
  LaunchBisection a =
      make_launch(ctx, P, desc, nrep, (uint32_t)first_move, num_moves, ctx->d_repids.as<int>(), ctx->d_reppos.as<int>(),
                  ctx->d_stats.as<int64_t>(), records ? ctx->d_records.as<pimc_move_record>() : nullptr,
                  ctx->d_queue.as<int>(), nullptr, ctx->d_bar.as<int>(), ctx->d_part.as<double>());
  if (P.table && (rc = table_launch_setup(ctx, P, nrep, precision, a))) return rc;
  if ((rc = launch_timed(ctx, P, a, precision, nrep))) return rc;
  if (!records && !stats) return PIMC_OK;   // asynchronous: counters via pimc_bisection_stats_take
  if (records)
    HIP_TRY(hipMemcpyAsync(records, ctx->d_records.p, (size_t)nrep * num_moves * sizeof(pimc_move_record),
                           hipMemcpyDeviceToHost, ctx->stream));
  std::vector<int64_t> st((size_t)nrep * 5);
  HIP_TRY(hipMemcpyAsync(st.data(), ctx->d_stats.p, st.size() * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if ((rc = take_abort(ctx))) return rc;
  if (stats) {
    float ms = 0.0f;
    const auto& ev = ctx->timed.back();
    HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
    stats->device_ms = ms;
    for (int r = 0; r < nrep; r++) {
      stats->moves += st[5 * r + 0];
      stats->accepted += st[5 * r + 1];
      stats->stages += st[5 * r + 2];
      stats->pair_evals += st[5 * r + 3];
      stats->k_evals += st[5 * r + 4];
    }
  }
  return PIMC_OK;
}

int pimc_bisection_set_plan(pimc_ctx* ctx, const pimc_bisection_plan* plan) {
  if (!ctx || !plan) return fail(PIMC_ERR_ARG, "bisection_set_plan: NULL argument");
  if (plan->block_threads != 256 && plan->block_threads != 512)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: block_threads must be 256 or 512");
  if (plan->link_split != 1 && plan->link_split != 2 && plan->link_split != 4)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: link_split must be 1, 2 or 4");
  if (plan->cluster < 0 || plan->cluster > kMaxCluster)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: cluster must be 0.." + std::to_string(kMaxCluster));
  if (plan->cluster > 1 && plan->persistent)
    return fail(PIMC_ERR_ARG, "bisection_set_plan: a cluster plan is not persistent");
  ctx->plan = *plan;
  ctx->plan.persistent = plan->persistent ? 1 : 0;
  ctx->plan.cluster = std::max(1, plan->cluster);
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
  BisectionPrep P0;
  // validate with a one-block plan: the cluster trials below are skipped where they cannot run
  const pimc_bisection_plan caller_plan = ctx->plan;
  ctx->plan.cluster = 1;
  int rc = prepare_bisection(ctx, replicas, nrep, desc, precision, P0);
  ctx->plan = caller_plan;
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = ctx->sync_descriptors())) return rc;
  if ((rc = ensure_acc(ctx))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  // save what trial moves change: both path copies, permutations, rho_k, the counters
  auto state_of = [&](int r) {
    HostReplica* H = ctx->rep(replicas[r]);
    return std::vector<DevBuf*>{&H->path0, &H->path1, &H->perm0, &H->perm1, &H->joinbuf, &H->rho0, &H->rho1};
  };
  std::vector<std::unique_ptr<DevBuf>> saved;
  for (int r = 0; r < nrep; r++)
    for (DevBuf* b : state_of(r)) {
      saved.push_back(std::make_unique<DevBuf>());
      if (b->bytes) {
        if ((rc = saved.back()->ensure(b->bytes))) return rc;
        HIP_TRY(hipMemcpyAsync(saved.back()->p, b->p, b->bytes, hipMemcpyDeviceToDevice, ctx->stream));
      }
    }
  std::vector<int64_t> acc((size_t)ctx->acc_cap * 5);
  HIP_TRY(hipMemcpyAsync(acc.data(), ctx->d_acc.p, acc.size() * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  auto restore = [&]() -> int {
    size_t i = 0;
    for (int r = 0; r < nrep; r++)
      for (DevBuf* b : state_of(r)) {
        const DevBuf& sv = *saved[i++];
        if (b->bytes) HIP_TRY(hipMemcpyAsync(b->p, sv.p, b->bytes, hipMemcpyDeviceToDevice, ctx->stream));
      }
    HIP_TRY(hipMemcpyAsync(ctx->d_acc.p, acc.data(), acc.size() * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PIMC_OK;
  };
  const pimc_bisection_plan keep = ctx->plan;
  pimc_bisection_plan best = keep;
  double best_ms = 1e300;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // every (block, link split) with one block per replica, a persistent grid, and
  // clusters of 2 .. kMaxCluster blocks per replica where all blocks fit at once; TABLE moves
  // run one 256-thread block per replica, so only their link split is tried
  std::vector<pimc_bisection_plan> plans;
  if (P0.table) {
    for (int ls : {1, 2, 4}) plans.push_back(pimc_bisection_plan{256, ls, 0, 1, 0.0});
  } else {
    for (int block : {256, 512})
      for (int ls : {1, 2, 4}) {
        for (int pers : {0, 1}) plans.push_back(pimc_bisection_plan{block, ls, pers, 1, 0.0});
        for (int c = 2; c <= kMaxCluster; c *= 2) plans.push_back(pimc_bisection_plan{block, ls, 0, c, 0.0});
      }
  }
  for (const pimc_bisection_plan& plan : plans) {
    ctx->plan = plan;
    BisectionPrep P;
    if ((rc = prepare_bisection(ctx, replicas, nrep, desc, precision, P))) {
      if (rc == PIMC_ERR_UNSUPPORTED && plan.cluster > 1) {   // a permuted replica: no cluster trial
        rc = PIMC_OK;
        continue;
      }
      break;
    }
    DevBuf ids, pos, st, q, bar, part;
    if ((rc = ids.ensure((size_t)nrep * 4)) || (rc = pos.ensure((size_t)nrep * 4)) ||
        (rc = st.ensure((size_t)nrep * 40)) || (rc = q.ensure(4)) || (rc = ensure_cluster_buffers(ctx, nrep, bar, part)))
      break;
    HIP_TRY(hipMemcpy(ids.p, P.ids.data(), (size_t)nrep * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pos.p, P.order.data(), (size_t)nrep * 4, hipMemcpyHostToDevice));
    LaunchBisection a =
        make_launch(ctx, P, desc, nrep, (uint32_t)first_move, trial_moves, ids.as<int>(), pos.as<int>(),
                    st.as<int64_t>(), nullptr, q.as<int>(), nullptr, bar.as<int>(), part.as<double>());
    if (P.table && (rc = table_launch_setup(ctx, P, nrep, precision, a))) break;
    if (plan.cluster > 1 && check_cluster_fits(ctx, P, a, precision, nrep) != PIMC_OK) continue;   // no trial
    const int grid = plan.persistent ? persistent_grid(ctx, P, a, precision, nrep) : nrep;
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    hipError_t e = P.table ? launch_bisection_table(P.nt, precision == PIMC_MIXED, P.lean, a, ctx->stream)
                           : launch_bisection(P.nt, precision == PIMC_MIXED, P.lean, a, plan.block_threads, grid,
                                              ctx->stream);
    if (e != hipSuccess) {
      rc = fail(PIMC_ERR_DEVICE, std::string("bisection_tune launch: ") + hipGetErrorString(e));
      break;
    }
    HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = take_abort(ctx))) break;
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = ctx->plan;
      best.trial_ms = ms;
    }
    if ((rc = restore())) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    ctx->plan = keep;
    (void)restore();
    return rc;
  }
  ctx->plan = best;
  if (out) *out = best;
  return PIMC_OK;
}

int pimc_sweep_graph_create(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, const pimc_bisection_desc* desc,
                            int64_t first_move, int32_t num_moves, int32_t precision, int32_t* graph_id) {
  if (!ctx || !desc || nrep <= 0 || !replicas || num_moves <= 0 || first_move < 0 || !graph_id ||
      first_move > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "sweep_graph_create: bad argument");
  BisectionPrep P;
  int rc = prepare_bisection(ctx, replicas, nrep, desc, precision, P);
  if (rc) return rc;
  if (P.table) return fail(PIMC_ERR_UNSUPPORTED, "sweep_graph_create: TABLE permutation moves are not captured in graphs");
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = ctx->sync_descriptors())) return rc;
  if ((rc = ensure_acc(ctx))) return rc;
  auto g = std::make_unique<SweepGraph>();
  if ((rc = g->ids.ensure((size_t)nrep * 4)) || (rc = g->pos.ensure((size_t)nrep * 4)) ||
      (rc = g->stats.ensure((size_t)nrep * 40)) || (rc = g->queue.ensure(4)) || (rc = g->first.ensure(4)) ||
      (rc = ensure_cluster_buffers(ctx, nrep, g->bar, g->part)))
    return rc;
  const uint32_t f0 = (uint32_t)first_move;
  HIP_TRY(hipMemcpy(g->ids.p, P.ids.data(), (size_t)nrep * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g->pos.p, P.order.data(), (size_t)nrep * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g->first.p, &f0, 4, hipMemcpyHostToDevice));
  const LaunchBisection a =
      make_launch(ctx, P, desc, nrep, 0u, num_moves, g->ids.as<int>(), g->pos.as<int>(), g->stats.as<int64_t>(),
                  nullptr, g->queue.as<int>(), g->first.as<uint32_t>(), g->bar.as<int>(), g->part.as<double>());
  if ((rc = check_cluster_fits(ctx, P, a, precision, nrep))) return rc;
  const int grid = ctx->plan.persistent ? persistent_grid(ctx, P, a, precision, nrep) : nrep;
  // one sweep: [queue reset] -> bisection moves -> first_move += num_moves; kernels only
  HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_bisection(P.nt, precision == PIMC_MIXED, P.lean, a, ctx->plan.block_threads, grid, ctx->stream);
  if (e == hipSuccess) e = launch_sweep_advance(g->first.as<uint32_t>(), num_moves, ctx->stream);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(PIMC_ERR_DEVICE, std::string("sweep_graph_create: capture failed: ") +
                                     hipGetErrorString(e != hipSuccess ? e : e2));
  }
  g->graph = graph;
  HIP_TRY(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
  g->num_moves = num_moves;
  g->first_move = first_move;
  g->reps = P.ids;
  ctx->graphs.push_back(std::move(g));
  *graph_id = (int32_t)ctx->graphs.size() - 1;
  return PIMC_OK;
}

int pimc_sweep_graph_launch(pimc_ctx* ctx, int32_t graph_id, int32_t num_sweeps) {
  if (!ctx || graph_id < 0 || graph_id >= (int)ctx->graphs.size() || !ctx->graphs[graph_id] || num_sweeps < 0)
    return fail(PIMC_ERR_ARG, "sweep_graph_launch: bad argument");
  SweepGraph& g = *ctx->graphs[graph_id];
  if (g.first_move + (int64_t)g.num_moves * (g.sweeps + num_sweeps) > (int64_t)0xffffffffLL)
    return fail(PIMC_ERR_ARG, "sweep_graph_launch: move numbers beyond 2^32");
  HIP_TRY(hipSetDevice(ctx->device));
  std::pair<hipEvent_t, hipEvent_t> ev;
  if (int rc_ev = timed_pair(ctx, ev)) return rc_ev;
  HIP_TRY(hipEventRecord(ev.first, ctx->stream));
  for (int r : g.reps)
    if (ctx->rep(r)->dev.core.lr_on) ctx->rep(r)->rho1_stale = true;   // replays update rho_k OLD only
  for (int i = 0; i < num_sweeps; i++) HIP_TRY(hipGraphLaunch(g.exec, ctx->stream));
  HIP_TRY(hipEventRecord(ev.second, ctx->stream));
  g.sweeps += num_sweeps;
  return PIMC_OK;
}

int pimc_sweep_graph_destroy(pimc_ctx* ctx, int32_t graph_id) {
  if (!ctx || graph_id < 0 || graph_id >= (int)ctx->graphs.size() || !ctx->graphs[graph_id])
    return fail(PIMC_ERR_ARG, "sweep_graph_destroy: bad argument");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->graphs[graph_id].reset();
  return PIMC_OK;
}

}  // extern "C"

namespace {

// What ActionsClass::Energy and the actions' d_dBeta need from the replica's tables (the
// fits' beta derivatives, dUlong, the long-range derivative tables) at `level`.
int check_dbeta_tables(pimc_ctx* ctx, HostReplica* H, int level, const char* who) {
  const RepCore& R = H->dev.core;
  for (int f = 0; f < R.ns * R.ns; f++) {
    const HostFit& HF = *ctx->fits[H->dev.sp.fit[f]];
    const FitDev& F = HF.dev;
    if ((F.kind == 0 && !F.e64) || (F.kind == 1 && !F.dbc))
      return fail(PIMC_ERR_STATE, std::string(who) + ": a pair fit has no beta-derivative table");
    if (level >= HF.num_levels()) return fail(PIMC_ERR_ARG, std::string(who) + ": level beyond the fit's tau levels");
    if (F.kind == 1 && R.use_long_range && (!((H->dulong_mask >> f) & 1u) || level >= R.ulong_levels))
      return fail(PIMC_ERR_STATE, std::string(who) + ": long-range pair fit without dUlong tables");
  }
  if (R.lr_on && (R.lr_mask & ~H->dlr_mask))
    return fail(PIMC_ERR_STATE, std::string(who) + ": long-range tables without beta derivatives");
  if (R.lr_on && level >= R.lr_levels) return fail(PIMC_ERR_ARG, std::string(who) + ": level beyond the long-range tables");
  return PIMC_OK;
}

// The actions' d_dBeta over links [slice1, slice2) at `level` of copy `mode` (the
// long-range sum over slices slice1 .. slice2), and the potential over the same slices:
// which = 1 short range, 2 kinetic, 4 long range (or together).  Fills en.du_short,
// en.kinetic, en.du_long, en.v_short and en.v_long of the parts asked for.
int dbeta_parts(pimc_ctx* ctx, int replica, int mode, int slice1, int slice2, int level, int which, pimc_energy& en) {
  HostReplica* H = ctx->rep(replica);
  const RepCore& R = H->dev.core;
  int nt = 0;
  bool v_known = true;   // every long-ranged pair has its Vlong tables (when long range is used)
  for (int f = 0; f < R.ns * R.ns; f++) {
    const HostFit& F = *ctx->fits[H->dev.sp.fit[f]];
    if (F.dev.kind == 0)
      nt = std::max(nt, F.num_terms);
    else if (R.use_long_range && !((R.vlong_mask >> f) & 1u))
      v_known = false;
  }
  for (int f = 0; f < R.ns * R.ns; f++) {
    const HostFit& F = *ctx->fits[H->dev.sp.fit[f]];
    if (F.dev.kind == 0 && F.num_terms != nt)
      return fail(PIMC_ERR_UNSUPPORTED, "d_dBeta: fits with different num_terms in one replica");
  }
  const bool vlong_known = !R.lr_on || !(R.lr_mask & ~R.vlong_mask);
  const int skip = 1 << level;
  const long long nlinks = (slice2 - slice1 + skip - 1) / skip;
  const long long npairs = (long long)R.N * (R.N - 1) / 2;
  const int pair_blocks = (which & 1) ? (int)std::max(1LL, std::min(2048LL, (npairs + kBlock - 1) / kBlock)) : 0;
  const int kin_blocks =
      (which & 2) ? (int)std::max(1LL, std::min(1024LL, ((long long)R.N * nlinks + kBlock - 1) / kBlock)) : 0;
  const int lr_slices = ((which & 4) && R.lr_on) ? (slice2 - slice1) / skip + 1 : 0;
  const size_t n_pp = 2 * (size_t)pair_blocks, n_kp = kin_blocks, n_lp = 6 * (size_t)lr_slices;
  int rc;
  if ((rc = ctx->d_out.ensure(std::max<size_t>(1, n_pp + n_kp + n_lp) * sizeof(double)))) return rc;
  if (lr_slices && (rc = ctx->d_in1.ensure((size_t)lr_slices * R.ns * R.nk * sizeof(d2v)))) return rc;
  double* dp = ctx->d_out.as<double>();
  if (lr_slices)   // LongRangeClass::d_dBeta recomputes rho_k from the positions (CalcRho_ks_Fast)
    launch_rho_slices(ctx->d_reps.as<ReplicaDev>(), replica, mode, slice1, skip, lr_slices, R.ns, R.nk,
                      2 * (R.maxk0 + R.maxk1 + R.maxk2) + 3, ctx->d_in1.as<d2v>(), ctx->stream);
  hipError_t e = launch_energy(nt, ctx->d_reps.as<ReplicaDev>(), ctx->d_fits.as<FitDev>(), replica, mode, slice1,
                               slice2, level, pair_blocks, kin_blocks, lr_slices, ctx->d_in1.as<d2v>(), dp, dp + n_pp,
                               dp + n_pp + n_kp, ctx->stream);
  if (e != hipSuccess) return fail(PIMC_ERR_DEVICE, std::string("d_dBeta launch: ") + hipGetErrorString(e));
  std::vector<double> h(n_pp + n_kp + n_lp);
  if (!h.empty()) HIP_TRY(hipMemcpyAsync(h.data(), dp, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (which & 1) {
    double v = 0.0;
    en.du_short = 0.0;
    for (int b = 0; b < pair_blocks; b++) {
      en.du_short += h[2 * b];
      v += h[2 * b + 1];
    }
    en.v_short = v_known ? v : nan;
  }
  if (which & 2) {
    en.kinetic = 0.0;
    for (int b = 0; b < kin_blocks; b++) en.kinetic += h[n_pp + b];
  }
  if (which & 4) {
    en.du_long = 0.0;
    en.v_long = 0.0;
    if (lr_slices) {
      double hh = 0.0, bg = 0.0, k0 = 0.0, vh = 0.0, vbg = 0.0, vk0 = 0.0;
      for (int sl = 0; sl < lr_slices; sl++) {
        const double* q = h.data() + n_pp + n_kp + 6 * sl;
        hh += q[0];
        bg += q[1];
        k0 += q[2];
        vh += q[3];
        vbg += q[4];
        vk0 += q[5];
      }
      en.du_long = hh + (R.use_background ? bg : k0);
      en.v_long = vh + (R.use_background ? vbg : vk0);
    }
    if (!vlong_known) en.v_long = nan;
  }
  return PIMC_OK;
}

}  // namespace

extern "C" {

int pimc_estimator_energy(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, int32_t mode, pimc_energy* out) {
  if (!ctx || nrep < 0 || (nrep > 0 && (!replicas || !out)) || mode < 0 || mode > 1)
    return fail(PIMC_ERR_ARG, "estimator_energy: bad argument");
  for (int r = 0; r < nrep; r++) {
    HostReplica* H = ctx->rep(replicas[r]);
    if (!H) return fail(PIMC_ERR_ARG, "estimator_energy: bad replica");
    int rc = check_dbeta_tables(ctx, H, 0, "estimator_energy");
    if (rc) return rc;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  for (int r = 0; r < nrep; r++) {
    pimc_energy en{};
    if ((rc = dbeta_parts(ctx, replicas[r], mode, 0, ctx->rep(replicas[r])->M, 0, 7, en))) return rc;
    en.total = en.kinetic + en.du_short + en.du_long;
    out[r] = en;
  }
  return PIMC_OK;
}

int pimc_action_d_dbeta(pimc_ctx* ctx, int32_t action, const pimc_range_req* reqs, int32_t nreq, double* out) {
  if (!ctx || nreq < 0 || (nreq > 0 && (!reqs || !out))) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad argument");
  if (action != PIMC_ACTION_SHORT_RANGE && action != PIMC_ACTION_KINETIC && action != PIMC_ACTION_LONG_RANGE)
    return fail(PIMC_ERR_ARG, "action_d_dbeta: bad action");
  for (int i = 0; i < nreq; i++) {
    const pimc_range_req& q = reqs[i];
    HostReplica* H = ctx->rep(q.replica);
    if (!H) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad replica id");
    if (q.mode != PIMC_OLDMODE && q.mode != PIMC_NEWMODE) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad mode");
    if (q.level < 0 || q.level >= PIMC_MAX_LEVELS) return fail(PIMC_ERR_ARG, "action_d_dbeta: bad level");
    const int skip = 1 << q.level;
    if (q.slice1 < 0 || q.slice1 > q.slice2 || q.slice1 + ((q.slice2 - q.slice1 + skip - 1) / skip) * skip > H->M)
      return fail(PIMC_ERR_ARG, "action_d_dbeta: bad slice range");
    if (action == PIMC_ACTION_LONG_RANGE && H->dev.core.nk == 0)
      return fail(PIMC_ERR_STATE, "action_d_dbeta: replica has no k set");
    if (action != PIMC_ACTION_KINETIC) {
      int rc = check_dbeta_tables(ctx, H, q.level, "action_d_dbeta");
      if (rc) return rc;
    }
  }
  if (nreq == 0) return PIMC_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  const int which = action == PIMC_ACTION_SHORT_RANGE ? 1 : action == PIMC_ACTION_KINETIC ? 2 : 4;
  for (int i = 0; i < nreq; i++) {
    pimc_energy en{};
    if ((rc = dbeta_parts(ctx, reqs[i].replica, reqs[i].mode, reqs[i].slice1, reqs[i].slice2, reqs[i].level, which,
                          en)))
      return rc;
    out[i] = which == 1 ? en.du_short : which == 2 ? en.kinetic : en.du_long;
  }
  return PIMC_OK;
}

int pimc_estimator_pair_correlation(pimc_ctx* ctx, int32_t replica, int32_t mode, int32_t s1, int32_t s2, double start,
                                    double end, int32_t num_points, int64_t* hist) {
  HostReplica* H = ctx ? ctx->rep(replica) : nullptr;
  if (!H || !hist || mode < 0 || mode > 1 || num_points < 2 || !(start < end))
    return fail(PIMC_ERR_ARG, "estimator_pair_correlation: bad argument");
  const RepCore& R = H->dev.core;
  if (s1 < 0 || s2 < 0 || s1 >= R.ns || s2 >= R.ns) return fail(PIMC_ERR_ARG, "estimator_pair_correlation: bad species");
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = ctx->sync_descriptors();
  if (rc) return rc;
  const int nbins = num_points - 1;
  if ((rc = ctx->d_img.ensure((size_t)nbins * sizeof(unsigned long long)))) return rc;
  HIP_TRY(hipMemsetAsync(ctx->d_img.p, 0, (size_t)nbins * sizeof(unsigned long long), ctx->stream));
  const RepSpecies& S = H->dev.sp;
  const int same = s1 == s2;
  const long long npairs = same ? (long long)S.count[s1] * (S.count[s1] - 1) / 2 : (long long)S.count[s1] * S.count[s2];
  const long long nitems = npairs * (R.nslices - 1);
  const int blocks = (int)std::max(1LL, std::min(4096LL, (nitems + kBlock - 1) / kBlock));
  const double delta = (end - start) / (double)(num_points - 1);
  hipError_t e = launch_pair_correlation(ctx->d_reps.as<ReplicaDev>(), replica, mode, S.first[s1], S.count[s1],
                                         S.first[s2], S.count[s2], same, start, 1.0 / delta, end, nbins,
                                         ctx->d_img.as<unsigned long long>(), nitems, blocks, ctx->stream);
  if (e != hipSuccess) return fail(PIMC_ERR_DEVICE, std::string("estimator_pair_correlation launch: ") + hipGetErrorString(e));
  std::vector<unsigned long long> h(nbins);
  HIP_TRY(hipMemcpyAsync(h.data(), ctx->d_img.p, (size_t)nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < nbins; i++) hist[i] += (int64_t)h[i];
  return PIMC_OK;
}

int pimc_bisection_stats_take(pimc_ctx* ctx, const int32_t* replicas, int32_t nrep, pimc_bisection_stats* stats) {
  if (!ctx || !stats || nrep < 0 || (nrep > 0 && !replicas)) return fail(PIMC_ERR_ARG, "bisection_stats_take: bad argument");
  for (int r = 0; r < nrep; r++)
    if (!ctx->rep(replicas[r])) return fail(PIMC_ERR_ARG, "bisection_stats_take: bad replica");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  memset(stats, 0, sizeof *stats);
  int rc = take_abort(ctx);
  if (rc) return rc;
  stats->device_ms = ctx->timed_ms;
  ctx->timed_ms = 0.0;
  for (auto& e : ctx->timed) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
    stats->device_ms += ms;
    ctx->spare.push_back(e);
  }
  ctx->timed.clear();
  if (ctx->acc_cap == 0) return PIMC_OK;
  std::vector<int64_t> acc((size_t)ctx->acc_cap * 5);
  HIP_TRY(hipMemcpy(acc.data(), ctx->d_acc.p, acc.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  for (int r = 0; r < nrep; r++) {
    int64_t* c = acc.data() + (size_t)replicas[r] * 5;
    stats->moves += c[0];
    stats->accepted += c[1];
    stats->stages += c[2];
    stats->pair_evals += c[3];
    stats->k_evals += c[4];
    std::fill(c, c + 5, 0);
  }
  HIP_TRY(hipMemcpy(ctx->d_acc.p, acc.data(), acc.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  return PIMC_OK;
}

}  // extern "C"
