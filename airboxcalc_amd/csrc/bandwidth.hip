// Practical HBM ceiling: a plain streaming copy of float4 (read + write), the
// yardstick the stream-collide kernel's achieved bandwidth is compared with.
#include <algorithm>

#include "abx_common.hpp"

namespace abx {

typedef float v4f __attribute__((ext_vector_type(4)));

// One 16-byte element per lane, non-temporal: the fastest plain copy measured
// on MI355X (about 6.6 TB/s).
__global__ __launch_bounds__(256) void k_copy4(const v4f* __restrict__ src, v4f* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace abx

using namespace abx;

extern "C" int abx_copy_bandwidth(abx_dev* dev, int64_t bytes, int iters, double* gbps, double* ms_per_copy) {
  if (!dev || bytes < 1024 || iters < 1 || !gbps) {
    set_error("abx_copy_bandwidth: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(dev->ordinal));
  const int64_t n = bytes / 16;
  v4f *a = nullptr, *b = nullptr;
  ABX_HIP(hipMalloc((void**)&a, (size_t)n * 16));
  if (hipMalloc((void**)&b, (size_t)n * 16) != hipSuccess) {
    (void)hipFree(a);
    set_error("abx_copy_bandwidth: hipMalloc failed");
    return -1;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemset(a, 0, (size_t)n * 16);
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e0, 0);
  for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(256), 0, 0, a, b, n);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const hipError_t err = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) {
    set_errorf("abx_copy_bandwidth: %s", hipGetErrorString(err));
    return -1;
  }
  *gbps = 2.0 * (double)n * 16 * iters / (ms * 1e-3) / 1e9;
  if (ms_per_copy) *ms_per_copy = ms / iters;
  return 0;
}

namespace abx {
namespace {

// The memory pattern of an AA even step with every position active: each
// lane reads the 19 population planes at its position and writes them back
// permuted (slot q -> slot opp q), in place, with non-temporal accesses as
// the stream-collide's even step makes them.  No collision; 152 B per
// position.  This is the ceiling of the stream-collide's access pattern, as
// k_copy4 is the ceiling of a plain copy: abx_aa_pattern_bandwidth reports
// the faster of a launch with as many workgroups per CU as fit and one
// capped at 4 (16 waves), the stream-collide's even-step default.
__global__ __launch_bounds__(256) void k_aa_pattern(float* __restrict__ f, int64_t P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  float v[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) v[q] = __builtin_nontemporal_load(f + q * P + i);
#pragma unroll
  for (int q = 0; q < 19; ++q) __builtin_nontemporal_store(v[q], f + OPP[q] * P + i);
}

}  // namespace
}  // namespace abx

extern "C" int abx_aa_pattern_bandwidth(abx_dev* dev, int64_t positions, int iters, double* gbps,
                                        double* ms_per_pass) {
  if (!dev || positions < 1024 || iters < 1 || !gbps) {
    set_error("abx_aa_pattern_bandwidth: bad argument");
    return -1;
  }
  ABX_HIP(hipSetDevice(dev->ordinal));
  float* f = nullptr;
  ABX_HIP(hipMalloc((void**)&f, (size_t)positions * 19 * 4));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipMemset(f, 0, (size_t)positions * 19 * 4);
  const unsigned grid = (unsigned)((positions + 255) / 256);
  float ms = 1e30f;
  for (const size_t lds : {(size_t)0, (size_t)(dev->lds_per_cu / 5 + 1024)}) {
    hipLaunchKernelGGL(k_aa_pattern, dim3(grid), dim3(256), lds, 0, f, positions);
    (void)hipEventRecord(e0, 0);
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_aa_pattern, dim3(grid), dim3(256), lds, 0, f, positions);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float m = 0;
    (void)hipEventElapsedTime(&m, e0, e1);
    ms = std::min(ms, m);
  }
  const hipError_t err = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(f);
  if (err != hipSuccess) {
    set_errorf("abx_aa_pattern_bandwidth: %s", hipGetErrorString(err));
    return -1;
  }
  *gbps = 152.0 * (double)positions * iters / (ms * 1e-3) / 1e9;
  if (ms_per_pass) *ms_per_pass = ms / iters;
  return 0;
}
