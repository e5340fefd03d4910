// Small device helpers shared by the HIP kernels of the PD step.
//
// All reductions here are deterministic: every workgroup writes one partial
// per launch and the consumer sums the partials in a fixed order in FP64.
// (The reference accumulates with float atomics, which is not reproducible:
// cloth_admm_helper.cu:146-161, inner_product.cu:14-45.)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace silk::amd::dev {

constexpr int kWave = 64;

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 ld3(const float* p, int i) {
  return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void st3(float* p, int i, f3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ f3 operator+(f3 a, f3 b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ f3 operator-(f3 a, f3 b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ f3 operator*(float s, f3 a) {
  return {s * a.x, s * a.y, s * a.z};
}
__device__ __forceinline__ float dot3(f3 a, f3 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

/// Sum of `v` over the workgroup in FP64, valid in thread 0.  blockDim.x must
/// be a multiple of 64 and at most 1024.  `lds` needs 16 doubles.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / kWave;
    for (int w = 0; w < nw; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

/// Sum of n doubles in fixed order by one workgroup; valid in all threads.
__device__ __forceinline__ double block_sum_array(const double* a, int n,
                                                  double* lds) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += a[i];
  double s = block_sum(v, lds);
  if (threadIdx.x == 0) lds[16] = s;
  __syncthreads();
  s = lds[16];
  __syncthreads();
  return s;
}

}  // namespace silk::amd::dev
