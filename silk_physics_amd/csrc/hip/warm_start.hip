// Learned CG warm start: feature assembly, BF16-MFMA inference and training
// kernels, the per-system judge and the data-set recorder (warm_start.hpp).
//
// The network is evaluated on tiles of 16 samples per wavefront with
// v_mfma_f32_16x16x16_bf16, samples along the MFMA's N dimension: the
// accumulator of a layer (lane l: sample l % 16, units 4 (l / 16) + i of its
// 16-unit block) is exactly the B operand the next layer needs (lane l:
// column l % 16, k = 4 (l / 16) + i), so activations stay in registers from
// input to output and back.  Weight gradients reduce over samples, which must
// sit in K: those operands are transposed through a per-wavefront LDS tile.
#include "hip/warm_start.hpp"

#include <algorithm>

#include "hip/device_common.hpp"

namespace silk::amd {

using namespace dev;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kClip = 8.0f;        // |prediction / sigma| and |target| bound
constexpr int kMlpThreads = 256;     // 4 wavefronts per workgroup
constexpr int kMlpMaxWaves = 1024;   // training pass

__device__ __forceinline__ unsigned short to_bf16(float v) { return float_to_half_bits(v, 2); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 ld_bf4(const unsigned short* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}

__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
  bf16x4 r;
  r[0] = static_cast<short>(to_bf16(v[0]));
  r[1] = static_cast<short>(to_bf16(v[1]));
  r[2] = static_cast<short>(to_bf16(v[2]));
  r[3] = static_cast<short>(to_bf16(v[3]));
  return r;
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 silu4(f32x4 z) {
  return {silu(z[0]), silu(z[1]), silu(z[2]), silu(z[3])};
}

// Orders the LDS accesses of one wavefront (a tile is written and read back
// by the same wavefront).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward operands of one lane: A fragments of the weights and the bias
// values in the accumulator layout.
struct Frags {
  bf16x4 w1[4];     // rows 16 mb + l%16 of W1, inputs 4 (l/16) + i
  bf16x4 w2[4][4];  // [mb][kb]: rows 16 mb + l%16, inputs 16 kb + 4 (l/16) + i
  bf16x4 w3[4];     // [kb]: row 0 (lanes l%16 == 0), inputs 16 kb + 4 (l/16) + i
  f32x4 b1[4], b2[4];
  float b3;
};

__device__ __forceinline__ void load_frags(const MlpDev& n, Frags& f) {
  const int l = threadIdx.x % kWave, r = l % 16, g = l / 16;
  const bf16x4 zero = {0, 0, 0, 0};
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f.w1[mb] = ld_bf4(n.w1 + (16 * mb + r) * kMlpIn + 4 * g);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      f.w2[mb][kb] = ld_bf4(n.w2 + (16 * mb + r) * kMlpHidden + 16 * kb + 4 * g);
    f.w3[mb] = r == 0 ? ld_bf4(n.w3 + 16 * mb + 4 * g) : zero;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      f.b1[mb][i] = n.params[kMlpB1 + 16 * mb + 4 * g + i];
      f.b2[mb][i] = n.params[kMlpB2 + 16 * mb + 4 * g + i];
    }
  }
  f.b3 = n.params[kMlpB3];
}

struct Fwd {
  f32x4 z1[4], z2[4];  // pre-activations
  bf16x4 h1[4], h2[4];
  float y;             // lanes 0-15: output of sample l
};

__device__ __forceinline__ void forward(const Frags& f, bf16x4 x, Fwd& o) {
  const int g = (threadIdx.x % kWave) / 16;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    o.z1[mb] = mfma(f.w1[mb], x, f.b1[mb]);
    o.h1[mb] = pack4(silu4(o.z1[mb]));
  }
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 acc = f.b2[mb];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(f.w2[mb][kb], o.h1[kb], acc);
    o.z2[mb] = acc;
    o.h2[mb] = pack4(silu4(acc));
  }
  f32x4 a3 = {g == 0 ? f.b3 : 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) a3 = mfma(f.w3[kb], o.h2[kb], a3);
  o.y = a3[0];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// ------------------------------------------------------------------ features

__global__ void __launch_bounds__(kBlock) k_warm_features(PcgDev p, WarmFeatArgs a,
                                                          unsigned short* feat,
                                                          float* sigma) {
  if (a.skip != nullptr && *a.skip != 0) return;
  const int b = blockIdx.x;
  if (static_cast<int>(threadIdx.x) >= p.blk_rows[b]) return;
  const int row = p.blk_row0[b] + threadIdx.x;
  const int sl = b * (kBlock / kSlice) + threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  const int off = p.slice_off[sl], width = p.slice_width[sl];
  const float di = a.d[row];
  float rn[3] = {0.0f, 0.0f, 0.0f}, dn[3] = {0.0f, 0.0f, 0.0f}, ksum = 0.0f;
  int cnt = 0;
  for (int k = 0; k < width; ++k) {
    const float val = p.ell_val[off + k * kSlice + lane];
    if (val == 0.0f) continue;  // padding (no coupling)
    const int col = p.ell_col[off + k * kSlice + lane];
    ksum += fabsf(val);
    ++cnt;
    const float inv_dj = 1.0f / a.d[col];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rn[c] += a.r[3 * col + c] * inv_dj;
      dn[c] += a.dxp[3 * col + c];
    }
  }
  const float inv_cnt = cnt > 0 ? 1.0f / static_cast<float>(cnt) : 0.0f;
  const float inv_di = 1.0f / di;
  const float f5 = a.cw != nullptr ? a.cw[row] * inv_di : 0.0f;
  const float f6 = a.pinned[row] ? 1.0f : 0.0f;
  const float f7 = ksum * inv_di;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float f[5] = {a.r[3 * row + c] * inv_di, rn[c] * inv_cnt, a.dxp[3 * row + c],
                  dn[c] * inv_cnt, a.h * a.v[3 * row + c]};
    float sg = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k) sg = fmaxf(sg, fabsf(f[k]));
    const float inv = sg > 0.0f ? 1.0f / sg : 0.0f;
    unsigned short h[16];
#pragma unroll
    for (int k = 0; k < 5; ++k) h[k] = to_bf16(f[k] * inv);
    h[5] = to_bf16(f5);
    h[6] = to_bf16(f6);
    h[7] = to_bf16(f7);
#pragma unroll
    for (int k = 8; k < 16; ++k) h[k] = 0;
    uint4* dst = reinterpret_cast<uint4*>(feat + (3 * static_cast<size_t>(row) + c) * kMlpIn);
    dst[0] = uint4{h[0] | (static_cast<unsigned>(h[1]) << 16), h[2] | (static_cast<unsigned>(h[3]) << 16),
                   h[4] | (static_cast<unsigned>(h[5]) << 16), h[6] | (static_cast<unsigned>(h[7]) << 16)};
    dst[1] = uint4{0u, 0u, 0u, 0u};
    sigma[3 * row + c] = sg;
  }
}

// ----------------------------------------------------------------- inference

__global__ void __launch_bounds__(kMlpThreads) k_mlp_predict(MlpDev net,
                                                             const unsigned short* feat,
                                                             const float* sigma, int64_t n,
                                                             float* x, float* delta,
                                                             const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  Frags f;
  load_frags(net, f);
  const int l = threadIdx.x % kWave, r = l % 16, g = l / 16;
  const bf16x4 zero = {0, 0, 0, 0};
  const int64_t ntile = (n + 15) / 16;
  const int64_t nw = static_cast<int64_t>(gridDim.x) * (kMlpThreads / kWave);
  for (int64_t t = blockIdx.x * (kMlpThreads / kWave) + threadIdx.x / kWave; t < ntile;
       t += nw) {
    const int64_t s = 16 * t + r;
    const bf16x4 xb = s < n ? ld_bf4(feat + s * kMlpIn + 4 * g) : zero;
    Fwd o;
    forward(f, xb, o);
    if (g == 0 && s < n) {
      const float dl = sigma[s] * fminf(fmaxf(o.y, -kClip), kClip);
      delta[s] = dl;
      if (x != nullptr) x[s] += dl;
    }
  }
}

// ----------------------------------------------------- projected start

// (A v)_i of one row for the three coordinates, difference form as the PCG's
// operator (the zero row sum of K is exact).
__device__ __forceinline__ void apply_row(const PcgDev& p, int b, int row,
                                          const float* __restrict__ v, float out[3]) {
  const int sl = b * (kBlock / kSlice) + threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  const int off = p.slice_off[sl], width = p.slice_width[sl];
  const float vi[3] = {v[3 * row], v[3 * row + 1], v[3 * row + 2]};
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < width; ++k) {
    const int col = p.ell_col[off + k * kSlice + lane];
    const float a = p.ell_val[off + k * kSlice + lane];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += a * (v[3 * col + c] - vi[c]);
  }
  const float di = p.d[row];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c] = acc[c] + di * vi[c];
}

// Per row block and coordinate c the sums the 2x2 projections need, with
// u1 = A v1, u2 = A v2 (v1 the prediction, v2 the previous correction; a
// null v is zero): [10 c + 0..2] v1.u1, v1.u2, v2.u2; [+3, 4] v1.r, v2.r;
// [+5..7] u1.u1, u1.u2, u2.u2; [+8, 9] u1.r, u2.r.
__global__ void __launch_bounds__(kBlock) k_warm_gram(PcgDev p, const float* v1,
                                                      const float* v2, const float* r,
                                                      double* part, const int* skip) {
  __shared__ double lds[17];
  if (skip != nullptr && *skip != 0) return;
  const int b = blockIdx.x;
  const int row = p.blk_row0[b] + threadIdx.x;
  const bool valid = static_cast<int>(threadIdx.x) < p.blk_rows[b];
  float u1[3] = {0.0f, 0.0f, 0.0f}, u2[3] = {0.0f, 0.0f, 0.0f};
  if (valid && v1 != nullptr) apply_row(p, b, row, v1, u1);
  if (valid && v2 != nullptr) apply_row(p, b, row, v2, u2);
  for (int c = 0; c < 3; ++c) {
    double t[10];
    for (int k = 0; k < 10; ++k) t[k] = 0.0;
    if (valid) {
      const double a1 = v1 != nullptr ? v1[3 * row + c] : 0.0;
      const double a2 = v2 != nullptr ? v2[3 * row + c] : 0.0;
      const double rr = r[3 * row + c];
      const double w1 = u1[c], w2 = u2[c];
      t[0] = a1 * w1; t[1] = a1 * w2; t[2] = a2 * w2;
      t[3] = a1 * rr; t[4] = a2 * rr;
      t[5] = w1 * w1; t[6] = w1 * w2; t[7] = w2 * w2;
      t[8] = w1 * rr; t[9] = w2 * rr;
    }
    for (int k = 0; k < 10; ++k) {
      const double sum = block_sum(t[k], lds);
      if (threadIdx.x == 0) part[30 * static_cast<size_t>(b) + 10 * c + k] = sum;
    }
  }
}

// One thread per (system, coordinate): the coefficients (c1, c2) of
// x0 + c1 v1 + c2 v2.  mode 0: the Galerkin projection (minimal A-norm
// error over the span); mode 1: the minimal residual ||r0 - c1 A v1 -
// c2 A v2||.  A direction whose Gram entry is not positive is dropped; a
// nearly dependent pair falls back to the better single direction.
__global__ void k_warm_coef(PcgDev p, const double* part, int mode, float* coef,
                            const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * p.nsys) return;
  const int sys = i / 3, c = i % 3;
  double t[10];
  for (int k = 0; k < 10; ++k) t[k] = 0.0;
  for (int q = p.sys_blk0[sys]; q < p.sys_blk0[sys + 1]; ++q)
    for (int k = 0; k < 10; ++k) t[k] += part[30 * static_cast<size_t>(q) + 10 * c + k];
  // G = [[g11, g12], [g12, g22]], rhs (h1, h2).
  double g11, g12, g22, h1, h2;
  if (mode == 0) {
    g11 = t[0]; g12 = t[1]; g22 = t[2]; h1 = t[3]; h2 = t[4];
  } else {
    g11 = t[5]; g12 = t[6]; g22 = t[7]; h1 = t[8]; h2 = t[9];
  }
  double c1 = 0.0, c2 = 0.0;
  const bool ok1 = g11 > 0.0 && isfinite(g11), ok2 = g22 > 0.0 && isfinite(g22);
  const double det = g11 * g22 - g12 * g12;
  if (ok1 && ok2 && det > 1e-10 * g11 * g22) {
    c1 = (h1 * g22 - h2 * g12) / det;
    c2 = (h2 * g11 - h1 * g12) / det;
  } else {
    // Single directions: the one that lowers the objective most.
    const double gain1 = ok1 ? h1 * h1 / g11 : 0.0, gain2 = ok2 ? h2 * h2 / g22 : 0.0;
    if (gain1 >= gain2 && ok1) c1 = h1 / g11;
    else if (ok2) c2 = h2 / g22;
  }
  if (!isfinite(c1) || !isfinite(c2)) c1 = c2 = 0.0;
  coef[6 * sys + c] = static_cast<float>(c1);
  coef[6 * sys + 3 + c] = static_cast<float>(c2);
}

__global__ void __launch_bounds__(kBlock) k_warm_apply(PcgDev p, const float* v1,
                                                       const float* v2, const float* coef,
                                                       float* x, const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  const int b = blockIdx.x;
  if (static_cast<int>(threadIdx.x) >= p.blk_rows[b]) return;
  const int row = p.blk_row0[b] + threadIdx.x;
  const float* cf = coef + 6 * p.blk_sys[b];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float v = x[3 * row + c];
    if (v1 != nullptr) v += cf[c] * v1[3 * row + c];
    if (v2 != nullptr) v += cf[3 + c] * v2[3 * row + c];
    x[3 * row + c] = v;
  }
}

// --------------------------------------------------------------- judge

__global__ void __launch_bounds__(kBlock) k_warm_judge(PcgDev p, const double* part2_prev,
                                                       const float* x_prev,
                                                       const float* r_prev, float* x,
                                                       float* r, int* counters,
                                                       const int* skip, double* r0d,
                                                       const double* r0d_prev) {
  __shared__ double lds[17];
  if (skip != nullptr && *skip != 0) return;
  const int b = blockIdx.x;
  const int sys = p.blk_sys[b];
  const int b0 = p.sys_blk0[sys], b1 = p.sys_blk0[sys + 1];
  double vn = 0.0, vo = 0.0;
  for (int q = b0 + threadIdx.x; q < b1; q += blockDim.x) {
    vn += p.part2[2 * q];
    vo += part2_prev[2 * q];
  }
  vn = block_sum(vn, lds);
  vo = block_sum(vo, lds);
  if (threadIdx.x == 0) lds[16] = vn < vo ? 1.0 : 0.0;
  __syncthreads();
  const bool keep = lds[16] != 0.0;
  if (!keep && static_cast<int>(threadIdx.x) < p.blk_rows[b]) {
    const int row = p.blk_row0[b] + threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[3 * row + c] = x_prev[3 * row + c];
      r[3 * row + c] = r_prev[3 * row + c];
      if (r0d != nullptr) r0d[3 * row + c] = r0d_prev[3 * row + c];
    }
  }
  if (threadIdx.x == 0 && b == b0) {
    atomicAdd(&counters[0], keep ? 1 : 0);
    atomicAdd(&counters[1], 1);
  }
}

// ------------------------------------------------------------------ record

__global__ void __launch_bounds__(kBlock) k_warm_record(
    const unsigned short* feat, const float* sigma, const float* dx, int64_t n,
    float rate, uint32_t salt, unsigned short* ds_feat, float* ds_target,
    unsigned long long* ds_count, int64_t cap, const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  const uint64_t thresh = static_cast<uint64_t>(static_cast<double>(rate) * 18446744073709551615.0);
  for (int64_t s = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; s < n;
       s += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (mix64(static_cast<uint64_t>(s) * 0x9E3779B97F4A7C15ULL + salt) >= thresh) continue;
    const float sg = sigma[s];
    if (!(sg > 0.0f)) continue;
    const unsigned long long idx = atomicAdd(ds_count, 1ULL);
    if (idx >= static_cast<unsigned long long>(cap)) continue;
    const uint4* src = reinterpret_cast<const uint4*>(feat + s * kMlpIn);
    uint4* dst = reinterpret_cast<uint4*>(ds_feat + idx * kMlpIn);
    dst[0] = src[0];
    dst[1] = src[1];
    ds_target[idx] = fminf(fmaxf(dx[s] / sg, -kClip), kClip);
  }
}

// ------------------------------------------------------------ weights (BF16)

__global__ void k_mlp_convert(const float* params, unsigned short* w1,
                              unsigned short* w2, unsigned short* w3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpHidden * kMlpIn) w1[i] = to_bf16(params[kMlpW1 + i]);
  if (i < kMlpHidden * kMlpHidden) w2[i] = to_bf16(params[kMlpW2 + i]);
  if (i < kMlpHidden) w3[i] = to_bf16(params[kMlpW3 + i]);
}

// ---------------------------------------------------------------- training

// Per-wavefront transposition tiles (BF16, [unit][sample]).
struct alignas(16) TrTile {
  unsigned short dz2[kMlpHidden][16];
  unsigned short h1[kMlpHidden][16];
  unsigned short dz1[kMlpHidden][16];
  unsigned short h2[kMlpHidden][16];
  unsigned short x[16][16];
  unsigned short dy[16];
};

// Sum over the 16 lanes of a row (same l / 16).
__device__ __forceinline__ float row16_sum(float v) {
  v += __shfl_xor(v, 1, kWave);
  v += __shfl_xor(v, 2, kWave);
  v += __shfl_xor(v, 4, kWave);
  v += __shfl_xor(v, 8, kWave);
  return v;
}

__global__ void __launch_bounds__(kMlpThreads) k_mlp_train(MlpDev net, MlpSamples data,
                                                           int64_t batch, int64_t step,
                                                           float* grad_part,
                                                           double* loss_part, float* y_out) {
  __shared__ TrTile tiles[kMlpThreads / kWave];
  Frags f;
  load_frags(net, f);
  const int l = threadIdx.x % kWave, r = l % 16, g = l / 16;
  const int wv = threadIdx.x / kWave;
  TrTile& T = tiles[wv];
  const bf16x4 zero = {0, 0, 0, 0};
  // Backward A fragments: W2^T (rows = inputs 16 mb + l%16, k = outputs
  // 16 kb + 4 (l/16) + i) and W3^T (rows = hidden 16 mb + l%16, k = output 0).
  bf16x4 w2t[4][4], w3t[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        w2t[mb][kb][i] = static_cast<short>(net.w2[(16 * kb + 4 * g + i) * kMlpHidden + 16 * mb + r]);
    w3t[mb] = zero;
    if (g == 0) w3t[mb][0] = static_cast<short>(net.w3[16 * mb + r]);
  }
  const f32x4 fz = {0.0f, 0.0f, 0.0f, 0.0f};
  f32x4 gw1[4], gw2[4][4], gw3[4], gb1[4], gb2[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    gw1[mb] = fz;
    gw3[mb] = fz;
    gb1[mb] = fz;
    gb2[mb] = fz;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) gw2[mb][nb] = fz;
  }
  float gb3 = 0.0f;
  double lsum = 0.0;
  const float inv_batch = 1.0f / static_cast<float>(batch);
  const int64_t ntile = (batch + 15) / 16;
  const int64_t wave = blockIdx.x * (kMlpThreads / kWave) + wv;
  const int64_t nw = static_cast<int64_t>(gridDim.x) * (kMlpThreads / kWave);
  for (int64_t t = wave; t < ntile; t += nw) {
    const int64_t i = 16 * t + r;
    const bool valid = i < batch;
    int64_t idx = i;
    if (step >= 0)
      idx = static_cast<int64_t>(mix64(static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ULL ^
                                       mix64(static_cast<uint64_t>(step) + 0x5117ULL)) %
                                 static_cast<uint64_t>(data.n));
    const bf16x4 xb = valid ? ld_bf4(data.feat + idx * kMlpIn + 4 * g) : zero;
    const float tgt = valid ? data.target[idx] : 0.0f;
    Fwd o;
    forward(f, xb, o);
    // Loss (lanes 0-15 hold the outputs): dL/dy = 2 (y - t) / batch.
    float dy = 0.0f;
    if (g == 0 && valid) {
      const float e = o.y - tgt;
      dy = 2.0f * e * inv_batch;
      lsum += static_cast<double>(e) * e;
      if (y_out != nullptr) y_out[i] = o.y;
    }
    gb3 += dy;
    bf16x4 dyb = zero;
    if (g == 0) dyb[0] = static_cast<short>(to_bf16(dy));
    // dZ2 = (W3^T dY) * silu'(Z2); dZ1 = (W2^T dZ2) * silu'(Z1).
    f32x4 dz2[4];
    bf16x4 dz2b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const f32x4 dh = mfma(w3t[mb], dyb, fz);
#pragma unroll
      for (int k = 0; k < 4; ++k) dz2[mb][k] = dh[k] * dsilu(o.z2[mb][k]);
      dz2b[mb] = pack4(dz2[mb]);
      gb2[mb] += dz2[mb];
    }
    f32x4 dz1[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = fz;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma(w2t[mb][kb], dz2b[kb], acc);
#pragma unroll
      for (int k = 0; k < 4; ++k) dz1[mb][k] = acc[k] * dsilu(o.z1[mb][k]);
      gb1[mb] += dz1[mb];
    }
    // Transpose the tile operands of the weight gradients through LDS:
    // lane l holds sample l % 16 and units 16 mb + 4 (l / 16) + k.
    wave_sync();
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int u = 16 * mb + 4 * g + k;
        T.dz2[u][r] = static_cast<unsigned short>(dz2b[mb][k]);
        T.h1[u][r] = static_cast<unsigned short>(o.h1[mb][k]);
        T.dz1[u][r] = to_bf16(dz1[mb][k]);
        T.h2[u][r] = static_cast<unsigned short>(o.h2[mb][k]);
      }
#pragma unroll
    for (int k = 0; k < 4; ++k) T.x[4 * g + k][r] = static_cast<unsigned short>(xb[k]);
    if (g == 0) T.dy[r] = static_cast<unsigned short>(dyb[0]);
    wave_sync();
    // dW2[out][in] += dZ2 H1^T; dW1[hidden][in] += dZ1 X^T; dW3[0][hidden] +=
    // dY H2^T.  A: row 16 mb + l%16, k = samples 4 (l/16) + i; B: column
    // 16 nb + l%16, k = samples 4 (l/16) + i.
    bf16x4 bh1[4], bh2[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      bh1[nb] = ld_bf4(&T.h1[16 * nb + r][4 * g]);
      bh2[nb] = ld_bf4(&T.h2[16 * nb + r][4 * g]);
    }
    const bf16x4 bx = ld_bf4(&T.x[r][4 * g]);
    const bf16x4 ady = r == 0 ? ld_bf4(&T.dy[4 * g]) : zero;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 adz2 = ld_bf4(&T.dz2[16 * mb + r][4 * g]);
      const bf16x4 adz1 = ld_bf4(&T.dz1[16 * mb + r][4 * g]);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) gw2[mb][nb] = mfma(adz2, bh1[nb], gw2[mb][nb]);
      gw1[mb] = mfma(adz1, bx, gw1[mb]);
      gw3[mb] = mfma(ady, bh2[mb], gw3[mb]);
    }
  }
  // Per-wavefront partials.  Accumulator lane l: column 16 nb + l%16, rows
  // 16 mb + 4 (l/16) + i.
  float* gp = grad_part + wave * static_cast<int64_t>(kMlpParams);
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int row = 16 * mb + 4 * g + k;
      gp[kMlpW1 + row * kMlpIn + r] = gw1[mb][k];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) gp[kMlpW2 + row * kMlpHidden + 16 * nb + r] = gw2[mb][nb][k];
      const float s1 = row16_sum(gb1[mb][k]);
      const float s2 = row16_sum(gb2[mb][k]);
      if (r == 0) {
        gp[kMlpB1 + row] = s1;
        gp[kMlpB2 + row] = s2;
      }
    }
  if (g == 0)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) gp[kMlpW3 + 16 * mb + r] = gw3[mb][0];
  const float s3 = row16_sum(gb3);
  if (l == 0) gp[kMlpB3] = s3;
  const double ls = sum64(lsum);
  if (l == 0) loss_part[wave] = ls / static_cast<double>(batch);
}

__global__ void __launch_bounds__(kBlock) k_mlp_reduce_adam(const float* grad_part,
                                                            const double* loss_part,
                                                            int waves, float* grad,
                                                            double* loss, float* params,
                                                            float* m, float* v, float lr,
                                                            int t) {
  const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx == 0) {
    double s = 0.0;
    for (int w = 0; w < waves; ++w) s += loss_part[w];
    *loss = s;
  }
  if (pidx >= kMlpParams) return;
  float gsum = 0.0f;
  for (int w = 0; w < waves; ++w) gsum += grad_part[static_cast<int64_t>(w) * kMlpParams + pidx];
  grad[pidx] = gsum;
  if (!(lr > 0.0f)) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mv = b1 * m[pidx] + (1.0f - b1) * gsum;
  const float vv = b2 * v[pidx] + (1.0f - b2) * gsum * gsum;
  m[pidx] = mv;
  v[pidx] = vv;
  const float mh = mv / (1.0f - powf(b1, static_cast<float>(t)));
  const float vh = vv / (1.0f - powf(b2, static_cast<float>(t)));
  params[pidx] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

void launch_warm_features(hipStream_t s, const PcgDev& p, const WarmFeatArgs& a,
                          unsigned short* feat, float* sigma) {
  if (p.nblk == 0) return;
  k_warm_features<<<p.nblk, kBlock, 0, s>>>(p, a, feat, sigma);
}

void launch_mlp_predict(hipStream_t s, const MlpDev& net, const unsigned short* feat,
                        const float* sigma, int64_t n, float* x, float* delta,
                        const int* skip) {
  if (n <= 0) return;
  const int64_t tiles = (n + 15) / 16;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (tiles + 3) / 4));
  k_mlp_predict<<<blocks, kMlpThreads, 0, s>>>(net, feat, sigma, n, x, delta, skip);
}

void launch_warm_project(hipStream_t s, const PcgDev& p, const float* v1, const float* v2,
                         const float* r, int mode, double* part, float* coef, float* x,
                         const int* skip) {
  if (p.nblk == 0) return;
  k_warm_gram<<<p.nblk, kBlock, 0, s>>>(p, v1, v2, r, part, skip);
  k_warm_coef<<<(3 * p.nsys + 63) / 64, 64, 0, s>>>(p, part, mode, coef, skip);
  k_warm_apply<<<p.nblk, kBlock, 0, s>>>(p, v1, v2, coef, x, skip);
}

void launch_warm_judge(hipStream_t s, const PcgDev& p, const double* part2_prev,
                       const float* x_prev, const float* r_prev, float* x,
                       float* r, int* counters, const int* skip, double* r0d,
                       const double* r0d_prev) {
  if (p.nblk == 0) return;
  if (r0d_prev == nullptr) r0d = nullptr;
  k_warm_judge<<<p.nblk, kBlock, 0, s>>>(p, part2_prev, x_prev, r_prev, x, r,
                                         counters, skip, r0d, r0d_prev);
}

void launch_warm_record(hipStream_t s, const unsigned short* feat,
                        const float* sigma, const float* dx, int64_t n, float rate,
                        uint32_t salt, unsigned short* ds_feat, float* ds_target,
                        unsigned long long* ds_count, int64_t cap, const int* skip) {
  if (n <= 0 || cap <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
  k_warm_record<<<blocks, kBlock, 0, s>>>(feat, sigma, dx, n, rate, salt, ds_feat,
                                          ds_target, ds_count, cap, skip);
}

void launch_mlp_convert(hipStream_t s, const float* params, unsigned short* w1,
                        unsigned short* w2, unsigned short* w3) {
  k_mlp_convert<<<(kMlpHidden * kMlpHidden + kBlock - 1) / kBlock, kBlock, 0, s>>>(
      params, w1, w2, w3);
}

int mlp_train_waves(int64_t batch) {
  const int64_t tiles = (batch + 15) / 16;
  const int64_t w = std::min<int64_t>(kMlpMaxWaves, tiles);
  return static_cast<int>((w + 3) / 4 * 4);
}

void launch_mlp_train_pass(hipStream_t s, const MlpDev& net, const MlpSamples& data,
                           int64_t batch, int64_t step, float* grad_part,
                           double* loss_part, int waves, float* y) {
  if (batch <= 0 || data.n <= 0 || waves <= 0) return;
  k_mlp_train<<<waves / 4, kMlpThreads, 0, s>>>(net, data, batch, step, grad_part,
                                                 loss_part, y);
}

void launch_mlp_reduce_adam(hipStream_t s, const float* grad_part,
                            const double* loss_part, int waves, float* grad,
                            double* loss, float* params, float* m, float* v,
                            float lr, int t) {
  k_mlp_reduce_adam<<<(kMlpParams + kBlock - 1) / kBlock, kBlock, 0, s>>>(
      grad_part, loss_part, waves, grad, loss, params, m, v, lr, t);
}

}  // namespace silk::amd
