// MAS-preconditioned conjugate gradient for the PD global step.
//
// The reference operator is A = k JWJ + kb CWC (x) I3 + diag(d) stored as a
// 3x3 BSR (cloth_assembly_l1_cache.cu:270-278, mas_cg_solver.cu:127-150).
// Every block of that matrix is a scalar times I3 and d is the same for the
// three coordinates of a vertex (mass, pin penalty, contact weight), so here
// the operator is stored once as a vertex-scalar sliced-ELL matrix and applied
// to the three coordinates together.  The MAS banks of the reference are
// 96x96 (32 vertices x 3); with the same structure they are 32x32 here and
// their inverse applies to the 32x3 bank residual.
//
// One PCG system per cloth, as in the reference (one MASCGSolver per cloth,
// cloth_admm_helper.cu:367-408).  All systems advance in the same launches;
// row blocks and MAS banks never straddle systems, and every scalar (rz, rr,
// pAp, alpha, beta, stopping) is per system.  Dot products are reduced in
// FP64 from per-block partials in a fixed order.
#include "hip/device_common.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

__device__ __forceinline__ double sys_sum(const PcgDev& p, int sys, int comp,
                                          double* lds) {
  const int b0 = p.sys_blk0[sys], b1 = p.sys_blk0[sys + 1];
  double v = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
    v += p.part[2 * b + comp];
  double s = block_sum(v, lds);
  if (threadIdx.x == 0) lds[16] = s;
  __syncthreads();
  s = lds[16];
  __syncthreads();
  return s;
}

// q = (K + D) src for the rows of block b; returns src_i . q_i per thread.
//
// K annihilates translations (J and C both map constants to zero:
// cloth_assembly_l2_cache.cpp:27-69,152-171), i.e. every row sums to zero.
// It is therefore applied in difference form, K src = sum_{j != i} K_ij
// (src_j - src_i), with only off-diagonal entries stored.  Rounding K to
// FP32 then cannot create a net force on rigidly translated positions; the
// naive form K_ii src_i + sum K_ij src_j does, and the spurious force grows
// with the absolute coordinates.
__device__ __forceinline__ double spmv_row(const PcgDev& p, int b,
                                           const float* __restrict__ src,
                                           float* __restrict__ out, int row,
                                           bool valid) {
  if (!valid) return 0.0;
  const int g = b * (kBlock / kSlice) + threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  const int off = p.slice_off[g];
  const int width = p.slice_width[g];
  const f3 s = ld3(src, row);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for (int k = 0; k < width; ++k) {
    const int c = p.ell_col[off + k * kSlice + lane];
    const float a = p.ell_val[off + k * kSlice + lane];
    ax += a * (src[3 * c] - s.x);
    ay += a * (src[3 * c + 1] - s.y);
    az += a * (src[3 * c + 2] - s.z);
  }
  const float dd = p.d[row];
  const f3 q = {ax + dd * s.x, ay + dd * s.y, az + dd * s.z};
  st3(out, row, q);
  return static_cast<double>(s.x) * q.x + static_cast<double>(s.y) * q.y +
         static_cast<double>(s.z) * q.z;
}

// ---- FP64 residual r = b - (K + D) x of the x-update system.  Also writes
// the float diagonal d used by the iterations and per-block partials
// (|r|^2, |b|^2) in part2.  r and b are formed in double from float inputs,
// so the cancellation between b and A x (both O(1), their difference small)
// does not lose the FP32 mantissa.
// Right-hand side and diagonal of one row of the x-update system in double.
__device__ __forceinline__ void rhs_row(const RhsArgs& ra, int row, double rhs[3],
                                        double& dd) {
  const ClothDev& c = ra.c;
  const double h = c.h;
  const double m = c.mass[row];
  dd = m / (h * h);
  for (int k = 0; k < 3; ++k) {
    // -m * inertia_mod, inertia_mod = -x/h^2 - v/h - a (admm_solver.cu:141-151)
    const double px = ra.prev_x[3 * row + k], pv = ra.prev_v[3 * row + k];
    rhs[k] = ra.elastic_rhs[3 * row + k] + c.bend_rhs[3 * row + k] +
             m * (px / (h * h) + pv / h + static_cast<double>(c.acc[k]));
  }
  if (c.pinned[row]) {
    dd += ra.pin_rho;
    for (int k = 0; k < 3; ++k)
      rhs[k] += static_cast<double>(ra.pin_rho) * c.pin_target[3 * row + k] -
                static_cast<double>(ra.pin_lambda[3 * row + k]);
  }
  if (ra.cw != nullptr) {
    const double w = ra.cw[row];
    if (w > 0.0) {
      dd += w;
      for (int k = 0; k < 3; ++k)
        rhs[k] += w * ra.cp[3 * row + k] + static_cast<double>(ra.ceta[3 * row + k]);
    }
  }
}

// K x of one row in double, difference form (see spmv_row).
__device__ __forceinline__ void kx_row(const PcgDev& p, int b, int row,
                                       const float* __restrict__ x, double ax[3]) {
  const int g = b * (kBlock / kSlice) + threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  const int off = p.slice_off[g];
  const int width = p.slice_width[g];
  const double xi[3] = {x[3 * row], x[3 * row + 1], x[3 * row + 2]};
  ax[0] = ax[1] = ax[2] = 0.0;
  for (int k = 0; k < width; ++k) {
    const int col = p.ell_col[off + k * kSlice + lane];
    const double a = p.ell_val[off + k * kSlice + lane];
    ax[0] += a * (static_cast<double>(x[3 * col]) - xi[0]);
    ax[1] += a * (static_cast<double>(x[3 * col + 1]) - xi[1]);
    ax[2] += a * (static_cast<double>(x[3 * col + 2]) - xi[2]);
  }
}

__global__ void __launch_bounds__(kBlock) k_residual(PcgDev p, RhsArgs ra,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ r_out,
                                                     int write_d) {
  __shared__ double lds[17];
  if (ra.skip != nullptr && *ra.skip != 0) return;
  const int b = blockIdx.x;
  const int row = p.blk_row0[b] + threadIdx.x;
  const bool valid = threadIdx.x < p.blk_rows[b];
  double rr = 0.0, bb = 0.0;
  if (valid) {
    double rhs[3], dd;
    rhs_row(ra, row, rhs, dd);
    if (write_d) p.d[row] = static_cast<float>(dd);
    // Difference form (see spmv_row): the zero row sum of K is exact.
    double ax[3];
    kx_row(p, b, row, x, ax);
    for (int k = 0; k < 3; ++k) {
      const double res = rhs[k] - (ax[k] + dd * x[3 * row + k]);
      r_out[3 * row + k] = static_cast<float>(res);
      if (ra.r0d != nullptr) ra.r0d[3 * row + k] = res;
      rr += res * res;
      bb += rhs[k] * rhs[k];
      if (ra.x_copy != nullptr) ra.x_copy[3 * row + k] = x[3 * row + k];
    }
  }
  const double s0 = block_sum(rr, lds);
  const double s1 = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    p.part2[2 * b] = s0;
    p.part2[2 * b + 1] = s1;
  }
}

// FP64 residuals of a finished solve, per row block: |b - A (x0 + dx)|^2
// (the iterate, x0 and dx applied separately so that their FP32 sum is never
// rounded), |b - A xc|^2 (the stored x) and |b|^2.
__global__ void __launch_bounds__(kBlock) k_solve_residuals(
    PcgDev p, RhsArgs ra, const float* __restrict__ x0,
    const float* __restrict__ dx, const float* __restrict__ xc,
    double* __restrict__ part3) {
  __shared__ double lds[17];
  if (ra.skip != nullptr && *ra.skip != 0) return;
  const int b = blockIdx.x;
  const int row = p.blk_row0[b] + threadIdx.x;
  const bool valid = threadIdx.x < p.blk_rows[b];
  double ri = 0.0, rc = 0.0, bb = 0.0;
  if (valid) {
    double rhs[3], dd;
    rhs_row(ra, row, rhs, dd);
    double a0[3], a1[3], a2[3];
    kx_row(p, b, row, x0, a0);
    kx_row(p, b, row, dx, a1);
    kx_row(p, b, row, xc, a2);
    for (int k = 0; k < 3; ++k) {
      const double xi = x0[3 * row + k], di = dx[3 * row + k], ci = xc[3 * row + k];
      const double e_it = rhs[k] - (a0[k] + dd * xi) - (a1[k] + dd * di);
      const double e_c = rhs[k] - (a2[k] + dd * ci);
      ri += e_it * e_it;
      rc += e_c * e_c;
      bb += rhs[k] * rhs[k];
    }
  }
  const double s0 = block_sum(ri, lds);
  const double s1 = block_sum(rc, lds);
  const double s2 = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    part3[3 * b] = s0;
    part3[3 * b + 1] = s1;
    part3[3 * b + 2] = s2;
  }
}

// Per system: the ratios of the block sums, folded into ctl's maxima.
__global__ void __launch_bounds__(kBlock) k_solve_residual_max(
    PcgDev p, const double* __restrict__ part3, const int* skip, AdmmCtl* ctl,
    int fold_iterate) {
  __shared__ double lds[17];
  if (skip != nullptr && *skip != 0) return;
  const int sys = blockIdx.x;
  const int b0 = p.sys_blk0[sys], b1 = p.sys_blk0[sys + 1];
  double s[3];
  for (int c = 0; c < 3; ++c) {
    double v = 0.0;
    for (int q = b0 + threadIdx.x; q < b1; q += blockDim.x) v += part3[3 * q + c];
    s[c] = block_sum(v, lds);
  }
  if (threadIdx.x == 0 && s[2] > 0.0) {
    const double it = s[0] / s[2], cm = s[1] / s[2];
    if (fold_iterate)
      atomicMax(&ctl->max_true_bits, static_cast<unsigned long long>(__double_as_longlong(it)));
    atomicMax(&ctl->max_commit_bits, static_cast<unsigned long long>(__double_as_longlong(cm)));
  }
}

// True-residual test of a finished solve in one launch.  Every row block
// adds (|b - A (x0 + dx)|^2, |b|^2) of its rows (x0 and dx applied
// separately, never summed in FP32) to part2; the last block to finish then
// tests every system against max(abs^2, rel^2 |b|^2), writes sys_done and
// ctl->verify_done, and folds the ratios into ctl->max_true_bits.
__global__ void __launch_bounds__(kBlock) k_iterate_verify(
    PcgDev p, RhsArgs ra, const float* __restrict__ x0, const float* __restrict__ dx,
    double* __restrict__ part2, int* sys_done, AdmmCtl* ctl, float rel_tol, int first,
    int final) {
  __shared__ double lds[17];
  __shared__ int last, all_ok;
  if (ctl->done != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->verify_done = 1;
    return;
  }
  const int b = blockIdx.x;
  const bool skip_blk = !first && sys_done[p.blk_sys[b]] != 0;
  double ri = 0.0, bb = 0.0;
  if (!skip_blk && threadIdx.x < p.blk_rows[b]) {
    const int row = p.blk_row0[b] + threadIdx.x;
    double rhs[3], dd;
    rhs_row(ra, row, rhs, dd);
    double a0[3], a1[3];
    kx_row(p, b, row, x0, a0);
    kx_row(p, b, row, dx, a1);
    for (int k = 0; k < 3; ++k) {
      const double e = rhs[k] - (a0[k] + dd * x0[3 * row + k]) -
                       (a1[k] + dd * dx[3 * row + k]);
      ri += e * e;
      bb += rhs[k] * rhs[k];
    }
  }
  const double s0 = block_sum(ri, lds);
  const double s1 = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    if (!skip_blk) {
      part2[2 * b] = s0;
      part2[2 * b + 1] = s1;
    }
    __threadfence();
    last = atomicAdd(&ctl->verify_arrive, 1) == static_cast<int>(gridDim.x) - 1;
    all_ok = 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const float rel = ctl->rel_tol > 0.0f ? ctl->rel_tol : rel_tol;
  const double rel2 = static_cast<double>(rel) * rel;
  const double abs2 = static_cast<double>(p.abs_tol) * p.abs_tol;
  for (int sys = 0; sys < p.nsys; ++sys) {
    if (!first && sys_done[sys] != 0) continue;  // same value in every thread
    const int b0 = p.sys_blk0[sys], b1 = p.sys_blk0[sys + 1];
    double v0 = 0.0, v1 = 0.0;
    for (int q = b0 + threadIdx.x; q < b1; q += blockDim.x) {
      v0 += __hip_atomic_load(&part2[2 * q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      v1 += __hip_atomic_load(&part2[2 * q + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const double rr = block_sum(v0, lds);
    const double bsq = block_sum(v1, lds);
    if (threadIdx.x == 0) {
      const bool ok = rr <= abs2 || rr <= rel2 * bsq;
      sys_done[sys] = ok ? 1 : 0;
      if (!ok) all_ok = 0;
      if ((ok || final) && bsq > 0.0)
        atomicMax(&ctl->max_true_bits,
                  static_cast<unsigned long long>(__double_as_longlong(rr / bsq)));
      if (!ok && final) ctl->pcg_unverified += 1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctl->verify_done = all_ok;
    ctl->verify_arrive = 0;
  }
}

__global__ void k_copy_unless(float* __restrict__ dst, const float* __restrict__ src,
                              int64_t n, const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

template <class T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_commit(int n3, float* __restrict__ x,
                         const float* __restrict__ dx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) x[i] += dx[i];
}

// ---- level-1 coarse sums of a 3-vector: cr[n] = sum_{rows in n} src
__global__ void k_coarse_sum(PcgDev p, const float* __restrict__ src,
                             int k_parity) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.n1node) return;
  (void)k_parity;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int e = p.cnode_ptr[n]; e < p.cnode_ptr[n + 1]; ++e) {
    const int row = p.cnode_list[e];
    sx += src[3 * row];
    sy += src[3 * row + 1];
    sz += src[3 * row + 2];
  }
  p.cr[3 * n] = sx;
  p.cr[3 * n + 1] = sy;
  p.cr[3 * n + 2] = sz;
}

// ---- level-1 coarse sums of the scalar diagonal
__global__ void k_coarse_dsum(PcgDev p, float* __restrict__ cd,
                              const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= p.n1node) return;
  float s = 0.0f;
  for (int e = p.cnode_ptr[n]; e < p.cnode_ptr[n + 1]; ++e)
    s += p.d[p.cnode_list[e]];
  cd[n] = s;
}

// Value of a bank slot: fine row, level-1 node or sum over level-1
// descendants of a deeper node.  Two threads per slot (t and t+32).
template <int NC>
__device__ __forceinline__ void load_bank_slot(const PcgDev& p, int bank,
                                               int level,
                                               const float* __restrict__ fine,
                                               const float* __restrict__ c1,
                                               float (*out)[NC],
                                               float (*tmp)[NC]) {
  const int slot = threadIdx.x & 31;
  const int half = threadIdx.x >> 5;
  const int m = p.bank_member[bank * kBankSlots + slot];
  float acc[NC];
  for (int c = 0; c < NC; ++c) acc[c] = 0.0f;
  if (m >= 0) {
    if (level == 0) {
      if (half == 0)
        for (int c = 0; c < NC; ++c) acc[c] = fine[NC * m + c];
    } else if (level == 1) {
      if (half == 0)
        for (int c = 0; c < NC; ++c) acc[c] = c1[NC * m + c];
    } else {
      const int e0 = p.cnode_ptr[m], e1 = p.cnode_ptr[m + 1];
      for (int e = e0 + half; e < e1; e += 2) {
        const int child = p.cnode_list[e];
        for (int c = 0; c < NC; ++c) acc[c] += c1[NC * child + c];
      }
    }
  }
  if (threadIdx.x < 64) {
    for (int c = 0; c < NC; ++c) tmp[threadIdx.x][c] = acc[c];
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    for (int c = 0; c < NC; ++c)
      out[slot][c] = tmp[slot][c] + tmp[slot + 32][c];
  }
  __syncthreads();
}

// ---- MAS factorization: one 64-thread workgroup per bank.
__global__ void __launch_bounds__(64) k_mas_factor(PcgDev p,
                                                   const float* __restrict__ cd,
                                                   const int* skip) {
  if (skip != nullptr && *skip != 0) return;
  const int bank = blockIdx.x;
  const int level = p.bank_level[bank];
  __shared__ float A[kBankSlots][kBankSlots + 1];
  __shared__ float dg[kBankSlots][1];
  __shared__ float tmp[64][1];
  const float* S = p.bank_static + static_cast<size_t>(bank) * 1024;
  for (int e = threadIdx.x; e < 1024; e += 64) A[e >> 5][e & 31] = S[e];
  load_bank_slot<1>(p, bank, level, p.d, cd, dg, tmp);
  if (threadIdx.x < 32) {
    const int s = threadIdx.x;
    float v = A[s][s] + dg[s][0];
    if (v == 0.0f) v = 1.0f;  // padding slot (mas_preconditioner.cu:730-745)
    A[s][s] = v;
  }
  __syncthreads();
  // In-place Gauss-Jordan inverse (SPD, no pivoting needed), in registers:
  // lane i of the first wavefront holds row i; the pivot row is read with
  // v_readlane (scalar registers) at each of the 32 fully unrolled steps.
  float row[kBankSlots];
  const int i = threadIdx.x & 31;
#pragma unroll
  for (int j = 0; j < kBankSlots; ++j) row[j] = A[i][j];
#pragma unroll
  for (int piv = 0; piv < kBankSlots; ++piv) {
    float prow[kBankSlots];
#pragma unroll
    for (int j = 0; j < kBankSlots; ++j)
      prow[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(row[j]), piv));
    const float ipv = 1.0f / prow[piv];
    const float aip = row[piv];
#pragma unroll
    for (int j = 0; j < kBankSlots; ++j) {
      float val;
      if (i == piv && j == piv) val = ipv;
      else if (i == piv) val = prow[j] * ipv;
      else if (j == piv) val = -aip * ipv;
      else val = row[j] - aip * (prow[j] * ipv);
      row[j] = val;
    }
  }
  __syncthreads();
  if (threadIdx.x < kBankSlots) {
#pragma unroll
    for (int j = 0; j < kBankSlots; ++j) A[i][j] = row[j];
  }
  __syncthreads();
  float* out = p.bank_inv + static_cast<size_t>(bank) * 1024;
  float amax = 0.0f;
  for (int e = threadIdx.x; e < 1024; e += 64) {
    // Symmetrize: the exact inverse is symmetric.
    const int i = e >> 5, j = e & 31;
    const float v = 0.5f * (A[i][j] + A[j][i]);
    out[e] = v;
    amax = fmaxf(amax, fabsf(v));
  }
  if (p.mas_half == 0) return;
  // 16-bit copy of the inverse (the PCG kernels read it instead): binary16
  // scaled by the bank's largest entry, bfloat16 unscaled.
  for (int off = kWave / 2; off > 0; off >>= 1)
    amax = fmaxf(amax, __shfl_xor(amax, off, kWave));
  const float scale = p.mas_half == 1 && amax > 0.0f ? amax : 1.0f;
  const float inv_scale = 1.0f / scale;
  if (threadIdx.x == 0) p.bank_scale[bank] = scale;
  unsigned short* oh = p.bank_inv_h + static_cast<size_t>(bank) * 1024;
  for (int e = threadIdx.x; e < 1024; e += 64)
    oh[e] = float_to_half_bits(out[e] * inv_scale, p.mas_half);
}

// ---- MAS bank solve: z_bank = inv(A_bank) r_bank (32 x 3).
__global__ void __launch_bounds__(64) k_bank_solve(PcgDev p,
                                                   const float* __restrict__ r,
                                                   const PcgState* __restrict__ st) {
  const int bank = blockIdx.x;
  if (st != nullptr && st[p.bank_sys[bank]].done) return;
  const int level = p.bank_level[bank];
  __shared__ float rb[kBankSlots][3];
  __shared__ float tmp[64][3];
  load_bank_slot<3>(p, bank, level, r, p.cr, rb, tmp);
  if (threadIdx.x >= 32) return;
  const int row = threadIdx.x;
  const float* inv = p.bank_inv + static_cast<size_t>(bank) * 1024;
  const unsigned short* invh = p.bank_inv_h + static_cast<size_t>(bank) * 1024;
  const float scale = p.mas_half != 0 ? p.bank_scale[bank] : 1.0f;
  float zx = 0.0f, zy = 0.0f, zz = 0.0f;
  for (int k = 0; k < kBankSlots; ++k) {
    const float a = p.mas_half != 0
                        ? scale * half_bits_to_float(invh[k * kBankSlots + row], p.mas_half)
                        : inv[k * kBankSlots + row];
    zx += a * rb[k][0];
    zy += a * rb[k][1];
    zz += a * rb[k][2];
  }
  const int m = p.bank_member[bank * kBankSlots + row];
  if (m < 0) return;
  float* dst = (level == 0) ? p.zl0 : p.cz;
  dst[3 * m] = zx;
  dst[3 * m + 1] = zy;
  dst[3 * m + 2] = zz;
}

// ---- z = sum over levels; partials (r.z, r.r)
__global__ void __launch_bounds__(kBlock) k_gather_z(PcgDev p,
                                                     const float* __restrict__ r,
                                                     const PcgState* __restrict__ st) {
  __shared__ double lds[17];
  const int b = blockIdx.x;
  if (st != nullptr && st[p.blk_sys[b]].done) return;
  const int row = p.blk_row0[b] + threadIdx.x;
  const bool valid = threadIdx.x < p.blk_rows[b];
  double rz = 0.0, rr = 0.0;
  if (valid) {
    f3 z = ld3(p.zl0, row);
    for (int l = 0; l < 4; ++l) {
      const int n = p.row_cnode[4 * row + l];
      if (n >= 0) z = z + ld3(p.cz, n);
    }
    st3(p.z, row, z);
    const f3 rv = ld3(r, row);
    rz = static_cast<double>(rv.x) * z.x + static_cast<double>(rv.y) * z.y +
         static_cast<double>(rv.z) * z.z;
    rr = static_cast<double>(rv.x) * rv.x + static_cast<double>(rv.y) * rv.y +
         static_cast<double>(rv.z) * rv.z;
  }
  const double s0 = block_sum(rz, lds);
  const double s1 = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    p.part[2 * b] = s0;
    p.part[2 * b + 1] = s1;
  }
}

// ---- init finalize: rz0, rr0, p = z
__global__ void __launch_bounds__(kBlock) k_init_final(PcgDev p, float rel_tol) {
  __shared__ double lds[17];
  const int b = blockIdx.x;
  const int sys = p.blk_sys[b];
  const double rz = sys_sum(p, sys, 0, lds);
  const double rr = sys_sum(p, sys, 1, lds);
  double bb = 0.0;
  {
    const int b0 = p.sys_blk0[sys], b1 = p.sys_blk0[sys + 1];
    double v = 0.0;
    for (int q = b0 + threadIdx.x; q < b1; q += blockDim.x) v += p.part2[2 * q + 1];
    bb = block_sum(v, lds);
  }
  const int row = p.blk_row0[b] + threadIdx.x;
  if (threadIdx.x < p.blk_rows[b]) {
    st3(p.p, row, ld3(p.z, row));
    st3(p.dx, row, f3{0.0f, 0.0f, 0.0f});
  }
  if (threadIdx.x == 0 && b == p.sys_blk0[sys]) {
    PcgState s;
    s.rz = rz;
    s.rr = rr;
    s.rr0 = rr;
    s.bb = bb;
    s.iter = 0;
    const bool ok = isfinite(rz);
    s.done = ok ? 0 : 1;
    s.status = ok ? -1 : 3;
    // Right-hand-side criterion: a warm start that already meets it takes no
    // iteration.
    const double rel2 = static_cast<double>(rel_tol) * rel_tol;
    const double abs2 = static_cast<double>(p.abs_tol) * p.abs_tol;
    if (ok && p.criterion >= 1 && (rr <= abs2 || rr <= rel2 * bb)) {
      s.done = 1;
      s.status = rr <= abs2 ? 0 : 1;
    }
    s.pad = 0;
    p.state[sys] = s;  // parity 0
  }
}

// ---- iteration phase 1: q = (K+D) p, partial p.q
__global__ void __launch_bounds__(kBlock) k_it_spmv(PcgDev p, int k) {
  __shared__ double lds[17];
  const int b = blockIdx.x;
  const PcgState* st = p.state + (k & 1) * p.nsys;
  if (st[p.blk_sys[b]].done) return;
  const int row = p.blk_row0[b] + threadIdx.x;
  const bool valid = threadIdx.x < p.blk_rows[b];
  const double v = spmv_row(p, b, p.p, p.q, row, valid);
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) p.part[2 * b] = s;
}

// ---- iteration phase 2: alpha, x += alpha p, r_out = r_in - alpha q
__global__ void __launch_bounds__(kBlock) k_it_update(PcgDev p, int k,
                                                      const float* __restrict__ r_in,
                                                      float* __restrict__ r_out) {
  __shared__ double lds[17];
  const int b = blockIdx.x;
  const int sys = p.blk_sys[b];
  const PcgState* st = p.state + (k & 1) * p.nsys;
  if (st[sys].done) return;
  const double pAp = sys_sum(p, sys, 0, lds);
  const double alpha_d = (fabs(pAp) < 1e-20) ? 0.0 : st[sys].rz / pAp;
  const float alpha = static_cast<float>(alpha_d);
  const int row = p.blk_row0[b] + threadIdx.x;
  if (threadIdx.x < p.blk_rows[b]) {
    const f3 pv = ld3(p.p, row);
    const f3 qv = ld3(p.q, row);
    st3(p.dx, row, ld3(p.dx, row) + alpha * pv);
    st3(r_out, row, ld3(r_in, row) - alpha * qv);
  }
}

// ---- iteration phase 4: stopping test, beta, p = z + beta p
__global__ void __launch_bounds__(kBlock) k_it_final(PcgDev p, int k,
                                                     float rel_tol) {
  __shared__ double lds[17];
  const int b = blockIdx.x;
  const int sys = p.blk_sys[b];
  const PcgState* st = p.state + (k & 1) * p.nsys;
  PcgState* nxt = p.state + ((k + 1) & 1) * p.nsys;
  const PcgState s = st[sys];
  if (s.done) {
    if (threadIdx.x == 0 && b == p.sys_blk0[sys]) nxt[sys] = s;
    return;
  }
  const double rz = sys_sum(p, sys, 0, lds);
  const double rr = sys_sum(p, sys, 1, lds);
  const int iter = s.iter + 1;
  const double abs2 = static_cast<double>(p.abs_tol) * p.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  int status = -1;
  if (iter >= p.max_iter) status = 2;
  else if (rr <= abs2) status = 0;
  else if (rr <= rel2 * (p.criterion >= 1 ? s.bb : s.rr0)) status = 1;
  const double beta_d = (fabs(s.rz) < 1e-20) ? 0.0 : rz / s.rz;
  const float beta = static_cast<float>(beta_d);
  const int row = p.blk_row0[b] + threadIdx.x;
  if (status < 0 && threadIdx.x < p.blk_rows[b])
    st3(p.p, row, ld3(p.z, row) + beta * ld3(p.p, row));
  if (threadIdx.x == 0 && b == p.sys_blk0[sys]) {
    PcgState n = s;
    n.rz = rz;
    n.rr = rr;
    n.iter = iter;
    n.status = status;
    n.done = status >= 0 ? 1 : 0;
    nxt[sys] = n;
  }
}

inline int grid_for(int n, int bs) { return (n + bs - 1) / bs; }

void mas_apply(hipStream_t s, const PcgDev& p, const float* r,
               const PcgState* st) {
  if (p.n1node > 0)
    k_coarse_sum<<<grid_for(p.n1node, kBlock), kBlock, 0, s>>>(p, r, 0);
  k_bank_solve<<<p.nbank, 64, 0, s>>>(p, r, st);
  k_gather_z<<<p.nblk, kBlock, 0, s>>>(p, r, st);
}

}  // namespace

void launch_mas_factor(hipStream_t s, const PcgDev& p, float* cd_scratch,
                       const int* skip) {
  if (p.nbank == 0) return;
  if (p.n1node > 0)
    k_coarse_dsum<<<grid_for(p.n1node, kBlock), kBlock, 0, s>>>(p, cd_scratch,
                                                                skip);
  k_mas_factor<<<p.nbank, 64, 0, s>>>(p, cd_scratch, skip);
}

void launch_pcg_residual(hipStream_t s, const PcgDev& p, const RhsArgs& ra) {
  if (p.nblk == 0) return;
  k_residual<<<p.nblk, kBlock, 0, s>>>(p, ra, p.x, p.r, 1);
}

void launch_pcg_start(hipStream_t s, const PcgDev& p, float rel_tol) {
  if (p.nblk == 0) return;
  mas_apply(s, p, p.r, nullptr);
  k_init_final<<<p.nblk, kBlock, 0, s>>>(p, rel_tol);
}

void launch_true_residual(hipStream_t s, const PcgDev& p, const RhsArgs& ra) {
  if (p.nblk == 0) return;
  k_residual<<<p.nblk, kBlock, 0, s>>>(p, ra, p.x, p.q, 0);
}

void launch_solve_residuals(hipStream_t s, const PcgDev& p, const RhsArgs& ra,
                            const float* x0, const float* dx, const float* xc,
                            double* part3, AdmmCtl* ctl, int fold_iterate) {
  if (p.nblk == 0) return;
  k_solve_residuals<<<p.nblk, kBlock, 0, s>>>(p, ra, x0, dx, xc, part3);
  k_solve_residual_max<<<p.nsys, kBlock, 0, s>>>(p, part3, ra.skip, ctl, fold_iterate);
}

void launch_iterate_verify(hipStream_t s, const PcgDev& p, const RhsArgs& ra,
                           const float* x0, const float* dx, double* part3,
                           int* sys_done, AdmmCtl* ctl, float rel_tol, int first,
                           int final) {
  if (p.nblk == 0) return;
  k_iterate_verify<<<p.nblk, kBlock, 0, s>>>(p, ra, x0, dx, part3, sys_done, ctl, rel_tol,
                                             first, final);
}

void launch_copy_unless(hipStream_t s, float* dst, const float* src, int64_t n,
                        const int* skip) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
  k_copy_unless<<<blocks, kBlock, 0, s>>>(dst, src, n, skip);
}

void launch_fill_u32(hipStream_t s, void* p, int64_t n, uint32_t v) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
  k_fill<uint32_t><<<blocks, kBlock, 0, s>>>(static_cast<uint32_t*>(p), n, v);
}

void launch_fill_u64(hipStream_t s, void* p, int64_t n, unsigned long long v) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(1024, (n + kBlock - 1) / kBlock));
  k_fill<unsigned long long>
      <<<blocks, kBlock, 0, s>>>(static_cast<unsigned long long*>(p), n, v);
}

void launch_pcg_commit(hipStream_t s, const PcgDev& p) {
  const int n3 = 3 * p.nv;
  if (n3 == 0) return;
  k_commit<<<grid_for(n3, kBlock), kBlock, 0, s>>>(n3, p.x, p.dx);
}

void launch_pcg_iters(hipStream_t s, const PcgDev& p, int k, int count,
                      float rel_tol) {
  if (p.nblk == 0) return;
  float* r_buf[2] = {p.r, p.r + 3 * static_cast<size_t>(p.nv)};
  for (int it = k; it < k + count; ++it) {
    const float* r_in = r_buf[it & 1];
    float* r_out = r_buf[(it + 1) & 1];
    const PcgState* st = p.state + (it & 1) * p.nsys;
    k_it_spmv<<<p.nblk, kBlock, 0, s>>>(p, it);
    k_it_update<<<p.nblk, kBlock, 0, s>>>(p, it, r_in, r_out);
    mas_apply(s, p, r_out, st);
    k_it_final<<<p.nblk, kBlock, 0, s>>>(p, it, rel_tol);
  }
}

void launch_mas_apply(hipStream_t s, const PcgDev& p) {
  mas_apply(s, p, p.r, nullptr);
}

}  // namespace silk::amd
