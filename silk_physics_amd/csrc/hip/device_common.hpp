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

// ---- cross-lane steps without LDS (DPP and the gfx950 permlane swaps).
// All lanes taking part must be active.

/// Value of lane l ^ 1 (DPP quad_perm [1, 0, 3, 2]).
constexpr int kDppXor1 = 0xB1;
/// Value of lane l ^ 2 (DPP quad_perm [2, 3, 0, 1]).
constexpr int kDppXor2 = 0x4E;
/// Value of lane 7 - l within each group of 8 (DPP row_half_mirror).
constexpr int kDppHalfMirror = 0x141;
/// Value of lane 15 - l within each row of 16 (DPP row_mirror).
constexpr int kDppMirror = 0x140;

/// Value of lane 16 * (l / 16) + n, for n = kDppRowBcast - 0x150 (DPP
/// row_newbcast:n; n in [0, 16)).
constexpr int kDppRowBcast = 0x150;

template <int kCtrl>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}

template <int kCtrl>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
      0, static_cast<int>(b), kCtrl, 0xF, 0xF, false));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_update_dpp(
      0, static_cast<int>(b >> 32), kCtrl, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}

/// Value of lane l ^ 16 (v_permlane16_swap: even and odd rows of 16 swap).
__device__ __forceinline__ unsigned xor16_u(unsigned v) {
  const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  return ((__lane_id() >> 4) & 1u) ? r[0] : r[1];
}

/// Value of lane l ^ 32 (v_permlane32_swap: the two halves swap).
__device__ __forceinline__ unsigned xor32_u(unsigned v) {
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return (__lane_id() >> 5) ? r[0] : r[1];
}

__device__ __forceinline__ float xor16_f(float v) {
  return __uint_as_float(xor16_u(__float_as_uint(v)));
}
__device__ __forceinline__ float xor32_f(float v) {
  return __uint_as_float(xor32_u(__float_as_uint(v)));
}
__device__ __forceinline__ double xor16_d(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = xor16_u(static_cast<unsigned>(b));
  const unsigned hi = xor16_u(static_cast<unsigned>(b >> 32));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ double xor32_d(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = xor32_u(static_cast<unsigned>(b));
  const unsigned hi = xor32_u(static_cast<unsigned>(b >> 32));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}

/// Sums over groups of 8, 32 or 64 lanes.  Every step adds two partial sums
/// of disjoint lane sets that are paired symmetrically (each lane of a pair
/// adds the same two values, in either order), so every lane of a group ends
/// with the same bits, and so does any group summing the same values in the
/// same lane positions.
template <typename T>
__device__ __forceinline__ T sum8(T v) {
  if constexpr (sizeof(T) == 8) {
    v += dpp_d<kDppXor1>(v);
    v += dpp_d<kDppXor2>(v);
    v += dpp_d<kDppHalfMirror>(v);
  } else {
    v += dpp_f<kDppXor1>(v);
    v += dpp_f<kDppXor2>(v);
    v += dpp_f<kDppHalfMirror>(v);
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T sum32(T v) {
  v = sum8(v);
  if constexpr (sizeof(T) == 8) {
    v += dpp_d<kDppMirror>(v);
    v += xor16_d(v);
  } else {
    v += dpp_f<kDppMirror>(v);
    v += xor16_f(v);
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T sum64(T v) {
  v = sum32(v);
  if constexpr (sizeof(T) == 8) v += xor32_d(v);
  else v += xor32_f(v);
  return v;
}

/// sum64 of two doubles with the steps of the two sums interleaved (the
/// same bits as two calls).
__device__ __forceinline__ void sum64_pair(double& a, double& b) {
  a += dpp_d<kDppXor1>(a);
  b += dpp_d<kDppXor1>(b);
  a += dpp_d<kDppXor2>(a);
  b += dpp_d<kDppXor2>(b);
  a += dpp_d<kDppHalfMirror>(a);
  b += dpp_d<kDppHalfMirror>(b);
  a += dpp_d<kDppMirror>(a);
  b += dpp_d<kDppMirror>(b);
  a += xor16_d(a);
  b += xor16_d(b);
  a += xor32_d(a);
  b += xor32_d(b);
}

// ---- 16-bit storage of MAS bank inverses.  kind 1: IEEE binary16 of the
// entry divided by the bank's scale (its largest magnitude), so that the
// entries sit in binary16's normal range whatever the stiffness; kind 2:
// bfloat16 (FP32 exponent range, scale 1).  Both round to nearest even.

__device__ __forceinline__ unsigned short float_to_half_bits(float v, int kind) {
  if (kind == 2) {
    unsigned u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
  }
  return __builtin_bit_cast(unsigned short, static_cast<_Float16>(v));
}

__device__ __forceinline__ float half_bits_to_float(unsigned short h, int kind) {
  if (kind == 2) return __uint_as_float(static_cast<unsigned>(h) << 16);
  return static_cast<float>(__builtin_bit_cast(_Float16, h));
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
