// pimc_calib.hip -- measured peaks for the roofline figures of the benchmark.
//
// Two microbenchmarks, timed with HIP events on a stream of their own:
//   HBM: a grid-stride copy of 16-byte elements (reads `bytes`, writes `bytes`);
//   LDS: every lane of a full grid reads 16-byte words from a 16 KB LDS buffer at
//        lane-consecutive (bank-conflict-free) addresses, four reads in flight.
// They are measurement utilities of the benchmark, not part of the action library.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kCopyBlock = 256;
constexpr int kLdsBlock = 256;
constexpr int kLdsWords = 1024;   // float4 words: 16 KB per block, eight blocks per CU fit

// Four 16-byte loads in flight per thread: element i + u * stride for u = 0..3.
__global__ void __launch_bounds__(kCopyBlock) k_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                     int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 4 * stride) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i + u * stride < n) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i + u * stride < n) dst[i + u * stride] = v[u];
  }
}

__global__ void __launch_bounds__(kLdsBlock) k_lds_read(int iters, float* __restrict__ out) {
  __shared__ float4 buf[kLdsWords];
  for (int i = threadIdx.x; i < kLdsWords; i += blockDim.x)
    buf[i] = make_float4((float)i, (float)(i + 1), (float)(i + 2), (float)(i + 3));
  __syncthreads();
  float4 a[4];
#pragma unroll
  for (int u = 0; u < 4; u++) a[u] = make_float4(0.f, 0.f, 0.f, 0.f);
  int idx = threadIdx.x;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const float4 v = buf[(idx + u * kLdsBlock) & (kLdsWords - 1)];
      a[u].x += v.x;
      a[u].y += v.y;
      a[u].z += v.z;
      a[u].w += v.w;
    }
    idx += 4 * kLdsBlock + 64;   // moves the window by one wave's worth of words each pass
  }
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < 4; u++) s += a[u].x + a[u].y + a[u].z + a[u].w;
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

struct Timer {
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  }
  bool init() {
    return hipStreamCreate(&s) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
           hipEventCreate(&e1) == hipSuccess;
  }
};

int num_cus(int* cus) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  return hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess ? 0 : 1;
}

}  // namespace

extern "C" {

// Copy bandwidth (bytes read + bytes written per second) of `bytes` (a multiple of 16)
// over `reps` timed copies after one warmup copy.  0 on success.
int pimc_calib_hbm(int64_t bytes, int reps, double* bytes_per_s) {
  if (bytes <= 0 || bytes % 16 != 0 || reps < 1 || !bytes_per_s) return 1;
  int cus = 0;
  if (num_cus(&cus)) return 2;
  Timer t;
  if (!t.init()) return 2;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return 3;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return 3;
  }
  const int64_t n = bytes / 16;
  const int grid = cus * 8;
  int rc = 0;
  if (hipMemsetAsync(a, 0, bytes, t.s) != hipSuccess) rc = 4;
  if (!rc) {
    k_copy<<<grid, kCopyBlock, 0, t.s>>>(a, b, n);
    (void)hipEventRecord(t.e0, t.s);
    for (int r = 0; r < reps; r++) k_copy<<<grid, kCopyBlock, 0, t.s>>>(a, b, n);
    (void)hipEventRecord(t.e1, t.s);
    if (hipEventSynchronize(t.e1) != hipSuccess || hipGetLastError() != hipSuccess) rc = 4;
  }
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, t.e0, t.e1) != hipSuccess) rc = 4;
  if (!rc) *bytes_per_s = 2.0 * (double)bytes * reps / (ms * 1e-3);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// LDS read bandwidth (bytes delivered to lanes per second) of a grid of 8 blocks of 256
// threads per CU, each lane reading 16 bytes per access, `iters` x 4 accesses per lane.
// *cus receives the CU count.  0 on success.
int pimc_calib_lds(int iters, double* bytes_per_s, int* cus) {
  if (iters < 1 || !bytes_per_s || !cus) return 1;
  if (num_cus(cus)) return 2;
  Timer t;
  if (!t.init()) return 2;
  const int grid = *cus * 8;
  float* out = nullptr;
  if (hipMalloc(&out, (size_t)grid * kLdsBlock * sizeof(float)) != hipSuccess) return 3;
  int rc = 0;
  k_lds_read<<<grid, kLdsBlock, 0, t.s>>>(iters, out);
  (void)hipEventRecord(t.e0, t.s);
  k_lds_read<<<grid, kLdsBlock, 0, t.s>>>(iters, out);
  (void)hipEventRecord(t.e1, t.s);
  if (hipEventSynchronize(t.e1) != hipSuccess || hipGetLastError() != hipSuccess) rc = 4;
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, t.e0, t.e1) != hipSuccess) rc = 4;
  if (!rc) *bytes_per_s = (double)grid * kLdsBlock * iters * 4.0 * 16.0 / (ms * 1e-3);
  (void)hipFree(out);
  return rc;
}

}  // extern "C"
