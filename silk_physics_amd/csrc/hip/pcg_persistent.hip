// Persistent MAS-preconditioned CG: one launch per solve.
//
// Same iteration as the reference MASCGSolver (mas_cg_solver.cu:193-311:
// alpha and beta guarded at |.| < 1e-20, stop on rr <= abs^2 or
// rr <= rel^2 rr0 or max_iter) and the same additive MAS preconditioner
// (mas_preconditioner.cu:601-728), without a launch per phase.
//
// Work split: a system (cloth) owns a group of workgroups; a workgroup owns
// up to eight consecutive 32-slot level-0 banks, one thread per slot.  A
// thread keeps its row's r, dx, z, p and q in registers for the whole solve;
// the workgroup keeps its eight bank inverses and every coarse bank inverse
// of the system in LDS.
//
// Per iteration k (p_k = z_k + beta_k p_{k-1}):
//   q = A p_k         neighbour values p_k(j) are formed on the fly from
//                     z_k(j) and p_{k-1}(j), both published before the last
//                     barrier, the same way their owner forms them
//   publish p.q partial and the per-level-1-node sums of q
//   B1                alpha; dx += alpha p; r -= alpha q; every workgroup
//                     updates its copy of the level-1 sums of r by
//                     cr -= alpha * qsum; level-0 banks solved from the
//                     workgroup's own rows; coarse banks solved from cr (all
//                     in LDS); z = z0 + coarse z; publish z, r.z, r.r
//   B2                beta and the stopping test (every workgroup sums the
//                     same partials in the same order: same decision)
//
// Cross-workgroup data (z, p, node sums, partials) is written and read with
// device-scope atomic stores and loads, which are coherent across the XCDs
// without invalidating the caches that hold the matrix; the barrier arrival
// is a release.  Barrier waits are bounded: a wait that gives up raises an
// abort flag, every workgroup leaves, and the host reports an error.
#include <cstdlib>

#include "hip/device_common.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

constexpr unsigned kSpinLimit = 1u << 22;
constexpr int kBanksPerWg = kPcgThreads / kBankSlots;  // 8

__device__ __forceinline__ void ast2(float* p, float a, float b) {
  const unsigned long long v =
      (static_cast<unsigned long long>(__float_as_uint(b)) << 32) |
      __float_as_uint(a);
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), v,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Store x, y, z of a 4-float record.
__device__ __forceinline__ void ast4(float* rec, f3 v) {
  ast2(rec, v.x, v.y);
  ast2(rec + 2, v.z, 0.0f);
}

__device__ __forceinline__ f3 ald4(const float* rec) {
  unsigned long long* q =
      reinterpret_cast<unsigned long long*>(const_cast<float*>(rec));
  const unsigned long long a =
      __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long b =
      __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return {__uint_as_float(static_cast<unsigned>(a)),
          __uint_as_float(static_cast<unsigned>(a >> 32)),
          __uint_as_float(static_cast<unsigned>(b))};
}

__device__ __forceinline__ void astd(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double aldd(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier of one workgroup group; count and gen sit in different
// 128-byte lines.  The arrival is a device-scope release; the wait polls
// with relaxed loads.  Returns false when the wait gave up.
__device__ bool group_barrier(unsigned* count, unsigned* gen, unsigned nwg,
                              unsigned& epoch, int* abort_flag, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned arrived = __hip_atomic_fetch_add(
        count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nwg - 1) {
      __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(gen, epoch + 1u, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == epoch) {
        if (++spins > kSpinLimit ||
            __hip_atomic_load(abort_flag, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    epoch += 1u;
    if (!ok)
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// Sum of partials a[4 * i + c], i < n, read with device-scope loads, in a
// fixed order; valid in every thread.
__device__ double wg_total(const double* a, int c, int n, double* lds) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += aldd(a + 4 * i + c);
  double s = block_sum(v, lds);
  if (threadIdx.x == 0) lds[16] = s;
  __syncthreads();
  s = lds[16];
  __syncthreads();
  return s;
}

__device__ __forceinline__ f3 fma3(float a, f3 x, f3 y) {
  return {fmaf(a, x.x, y.x), fmaf(a, x.y, y.y), fmaf(a, x.z, y.z)};
}

__device__ __forceinline__ double ddot3(f3 a, f3 b) {
  return static_cast<double>(a.x) * b.x + static_cast<double>(a.y) * b.y +
         static_cast<double>(a.z) * b.z;
}

__device__ __forceinline__ f3 lds4(const float* base, int i) {
  return {base[4 * i], base[4 * i + 1], base[4 * i + 2]};
}

// Sum over the 32 lanes of a half-wave (fixed butterfly order).
__device__ __forceinline__ f3 half_sum(f3 v) {
  for (int m = kBankSlots / 2; m > 0; m >>= 1) {
    v.x += __shfl_xor(v.x, m, kBankSlots);
    v.y += __shfl_xor(v.y, m, kBankSlots);
    v.z += __shfl_xor(v.z, m, kBankSlots);
  }
  return v;
}

// Workgroup sums of two doubles at once, valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* lds) {
  for (int off = kWave / 2; off > 0; off >>= 1) {
    a += __shfl_down(a, off, kWave);
    b += __shfl_down(b, off, kWave);
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) {
    lds[wid] = a;
    lds[8 + wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) {
      a += lds[w];
      b += lds[8 + w];
    }
  }
  __syncthreads();
}

// Totals of partials a[4 i + c0] and a[4 i + c1], i < n, read with
// device-scope loads, summed in a fixed order; valid in every thread.
__device__ void wg_total2(const double* a, int c0, int c1, int n, double* lds,
                          double& t0, double& t1) {
  double v0 = 0.0, v1 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    v0 += aldd(a + 4 * i + c0);
    v1 += aldd(a + 4 * i + c1);
  }
  block_sum2(v0, v1, lds);
  if (threadIdx.x == 0) {
    lds[16] = v0;
    lds[15] = v1;
  }
  __syncthreads();
  t0 = lds[16];
  t1 = lds[15];
  __syncthreads();
}

__global__ void __launch_bounds__(kPcgThreads)
    k_pcg_persistent(PersistDev a, float rel_tol) {
  // Dynamic LDS: [8 level-0 inverses][ncb_max coarse inverses]
  //              [n1_max level-1 sums][ncn_max coarse z][ncn_max coarse sums]
  //              (4 floats per node), then the coarse bank members, the
  //              descendant-list offsets and the descendant lists (ints).
  extern __shared__ float dyn[];
  __shared__ float rb[kPcgThreads][3];
  __shared__ float cbuf[kPcgThreads][3];
  __shared__ double red[17];
  __shared__ int s_ok;

  int sys = -1;
  for (int s = 0; s < a.nsys; ++s)
    if (static_cast<int>(blockIdx.x) >= a.sys[s].wg0 &&
        static_cast<int>(blockIdx.x) < a.sys[s].wg0 + a.sys[s].nwg)
      sys = s;
  if (sys < 0) return;  // spare workgroup: takes no part in any barrier
  // Stopped ADMM loop: nothing to solve.  ctl->done is written only by
  // other kernels, so every workgroup reads the same value.
  if (a.ctl != nullptr && a.ctl->done != 0) return;
  if (a.ctl != nullptr) rel_tol = a.ctl->rel_tol;
  const PcgSysDev S = a.sys[sys];
  unsigned* bcount = a.bar + kPcgBarrierStride * sys;
  unsigned* bgen = bcount + kPcgBarrierStride / 2;
  unsigned epoch = 0;
  if (threadIdx.x == 0)
    epoch = __hip_atomic_load(bgen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned W = static_cast<unsigned>(S.nwg);
  const int j = static_cast<int>(blockIdx.x) - S.wg0;
  const int nb0 = S.nslot / kBankSlots;
  const int b_lo = (j * nb0) / S.nwg, b_hi = ((j + 1) * nb0) / S.nwg;
  const int lb = b_lo + static_cast<int>(threadIdx.x) / kBankSlots;
  const int hw = threadIdx.x / kBankSlots;
  const int s_in_bank = threadIdx.x & (kBankSlots - 1);
  const int base = threadIdx.x & ~(kBankSlots - 1);
  const bool in_bank = lb < b_hi;
  const int slot = S.slot0 + kBankSlots * lb + s_in_bank;
  const int row = in_bank ? a.slot_row[slot] : -1;
  const bool own = row >= 0;
  const bool coarse = S.ncb > 0;
  const size_t ns4 = 4 * static_cast<size_t>(a.nslot_total);

  float* inv0 = dyn;                                           // 8 x 1024
  float* invc = inv0 + kBanksPerWg * 1024;                     // ncb_max x 1024
  float* crl = invc + static_cast<size_t>(a.ncb_max) * 1024;  // 4 x n1_max
  float* czl = crl + 4 * static_cast<size_t>(a.n1_max);       // 4 x ncn_max
  float* csum = czl + 4 * static_cast<size_t>(a.ncn_max);     // 4 x ncn_max
  int* cbm = reinterpret_cast<int*>(csum + 4 * static_cast<size_t>(a.ncn_max));
  int* cnp = cbm + kBankSlots * a.ncb_max;                     // ncn_max + 1
  int* cnl = cnp + a.ncn_max + 1;                              // cnl_max

  // Static data into LDS.
  for (int e = threadIdx.x; e < kBanksPerWg * 1024; e += kPcgThreads) {
    const int b = b_lo + e / 1024;
    inv0[e] = b < b_hi
                  ? a.bank_inv[static_cast<size_t>(S.bank0 + b) * 1024 + (e & 1023)]
                  : 0.0f;
  }
  for (int e = threadIdx.x; e < S.ncb * 1024; e += kPcgThreads)
    invc[e] = a.bank_inv[static_cast<size_t>(a.cb_bank[S.cb0 + e / 1024]) * 1024 +
                         (e & 1023)];
  for (int e = threadIdx.x; e < S.ncb * kBankSlots; e += kPcgThreads)
    cbm[e] = a.cb_member[kBankSlots * S.cb0 + e];
  const int cn_first = S.ncn > 0 ? a.cn_ptr[S.cn0] : 0;
  for (int e = threadIdx.x; e <= S.ncn; e += kPcgThreads)
    cnp[e] = a.cn_ptr[S.cn0 + e] - cn_first;
  for (int e = threadIdx.x; S.ncn > 0 && e < a.cn_ptr[S.cn0 + S.ncn] - cn_first;
       e += kPcgThreads)
    cnl[e] = a.cn_list[cn_first + e];

  int e_off = 0, e_w = 0, e_lane = 0;
  if (own) {
    const int sl = slot / kSlice;
    e_off = a.sell_off[sl];
    e_w = a.sell_width[sl];
    e_lane = slot % kSlice;
  }
  // This row's neighbours (slot-ELL, padded to whole chunks with the row's
  // own slot and weight 0) and coarse nodes, kept in registers.
  const bool reg_ell = e_w <= kPcgNbr;
  int ncol[kPcgNbr];
  float nval[kPcgNbr];
#pragma unroll
  for (int k = 0; k < kPcgNbr; ++k) {
    const bool v = own && reg_ell && k < e_w;
    ncol[k] = v ? a.sell_col[e_off + k * kSlice + e_lane] : slot;
    nval[k] = v ? a.sell_val[e_off + k * kSlice + e_lane] : 0.0f;
  }
  int scn[4] = {-1, -1, -1, -1};
  if (in_bank && coarse)
    for (int l = 0; l < 4; ++l) scn[l] = a.slot_cn[4 * slot + l];
  const float dd = own ? a.d[row] : 0.0f;
  f3 r = own ? ld3(a.r0, row) : f3{0.0f, 0.0f, 0.0f};
  f3 dx = {0.0f, 0.0f, 0.0f};
  f3 p = {0.0f, 0.0f, 0.0f};
  // Level-1 nodes of a bank have consecutive ids: first and count per bank.
  int n1_first = scn[0] >= 0 ? scn[0] : 0x7fffffff;
  int n1_last = scn[0];
  for (int m = kBankSlots / 2; m > 0; m >>= 1) {
    n1_first = min(n1_first, __shfl_xor(n1_first, m, kBankSlots));
    n1_last = max(n1_last, __shfl_xor(n1_last, m, kBankSlots));
  }
  const int n1_count = n1_last >= 0 ? n1_last - n1_first + 1 : 0;

  double* my_part = a.part + 4 * blockIdx.x;
  const double* grp_part = a.part + 4 * S.wg0;
  __syncthreads();

  // Sum of v over the slots of this thread's level-1 node, published by the
  // first slot of the node (level-1 nodes never straddle banks).
  auto publish_node_sums = [&](f3 v) {
    for (int k = 0; k < n1_count; ++k) {
      const int target = n1_first + k;
      const f3 t = half_sum(scn[0] == target ? v : f3{0.0f, 0.0f, 0.0f});
      if (s_in_bank == 0) ast4(a.nsum + 4 * static_cast<size_t>(S.cn0 + target), t);
    }
  };

  // z = M r from this thread's r (registers), the workgroup's banks and the
  // level-1 sums in crl.
  auto mas = [&](f3 rv) -> f3 {
    rb[threadIdx.x][0] = rv.x;
    rb[threadIdx.x][1] = rv.y;
    rb[threadIdx.x][2] = rv.z;
    __syncthreads();
    f3 z0 = {0.0f, 0.0f, 0.0f};
    if (in_bank) {
      const float* inv = inv0 + hw * 1024;
      for (int k = 0; k < kBankSlots; ++k) {
        const float w = inv[k * kBankSlots + s_in_bank];
        z0.x += w * rb[base + k][0];
        z0.y += w * rb[base + k][1];
        z0.z += w * rb[base + k][2];
      }
    }
    if (!coarse) return z0;
    // Residual sums of the coarse nodes, level by level from the level-1
    // sums: one half-wave per node, children strided over its lanes.
    for (int n = threadIdx.x; n < S.n1; n += kPcgThreads) {
      csum[4 * n] = crl[4 * n];
      csum[4 * n + 1] = crl[4 * n + 1];
      csum[4 * n + 2] = crl[4 * n + 2];
    }
    __syncthreads();
    for (int l = 2; l <= 4; ++l) {
      const int lo = S.lvl[l - 1], hi = S.lvl[l];
      if (lo >= hi) break;
      for (int n = lo + hw; n < hi; n += kBanksPerWg) {
        f3 v = {0.0f, 0.0f, 0.0f};
        for (int e = cnp[n] + s_in_bank; e < cnp[n + 1]; e += kBankSlots)
          v = v + lds4(csum, cnl[e]);
        v = half_sum(v);
        if (s_in_bank == 0) {
          csum[4 * n] = v.x;
          csum[4 * n + 1] = v.y;
          csum[4 * n + 2] = v.z;
        }
      }
      __syncthreads();
    }
    for (int round = 0; round < S.ncb; round += kBanksPerWg) {
      const int cb = round + hw;
      int m = -1;
      f3 v = {0.0f, 0.0f, 0.0f};
      if (cb < S.ncb) {
        m = cbm[kBankSlots * cb + s_in_bank];
        if (m >= 0) v = lds4(csum, m);
      }
      cbuf[threadIdx.x][0] = v.x;
      cbuf[threadIdx.x][1] = v.y;
      cbuf[threadIdx.x][2] = v.z;
      __syncthreads();
      if (cb < S.ncb && m >= 0) {
        const float* inv = invc + static_cast<size_t>(cb) * 1024;
        f3 zc = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < kBankSlots; ++k) {
          const float w = inv[k * kBankSlots + s_in_bank];
          zc.x += w * cbuf[base + k][0];
          zc.y += w * cbuf[base + k][1];
          zc.z += w * cbuf[base + k][2];
        }
        czl[4 * m] = zc.x;
        czl[4 * m + 1] = zc.y;
        czl[4 * m + 2] = zc.z;
      }
      __syncthreads();
    }
    if (in_bank) {
#pragma unroll
      for (int l = 0; l < 4; ++l)
        if (scn[l] >= 0) z0 = z0 + lds4(czl, scn[l]);
    }
    return z0;
  };

  // ---- prologue: level-1 sums of r0, z0 = M r0, r.z, r.r
  if (coarse) {
    publish_node_sums(r);
    if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
    for (int n = threadIdx.x; n < S.n1; n += kPcgThreads) {
      const f3 v = ald4(a.nsum + 4 * static_cast<size_t>(S.cn0 + n));
      crl[4 * n] = v.x;
      crl[4 * n + 1] = v.y;
      crl[4 * n + 2] = v.z;
    }
    __syncthreads();
  }
  f3 z = mas(r);
  if (own) ast4(a.zbuf + 4 * static_cast<size_t>(slot), z);
  {
    double s0 = own ? ddot3(r, z) : 0.0, s1 = own ? ddot3(r, r) : 0.0;
    block_sum2(s0, s1, red);
    if (threadIdx.x == 0) {
      astd(my_part + 1, s0);
      astd(my_part + 2, s1);
    }
  }
  if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
  double rz, rr;
  wg_total2(grp_part, 1, 2, S.nwg, red, rz, rr);
  const double rr0 = rr;
  const double abs2 = static_cast<double>(a.abs_tol) * a.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  int status = isfinite(rz) ? -1 : 3;
  int iter = 0;
  float beta = 0.0f;

  while (status < 0) {
    const float* p_prev = a.pbuf + static_cast<size_t>((iter + 1) & 1) * ns4;
    float* p_cur = a.pbuf + static_cast<size_t>(iter & 1) * ns4;
    p = (iter == 0) ? z : fma3(beta, p, z);
    if (own) ast4(p_cur + 4 * static_cast<size_t>(slot), p);
    // q = (K + D) p in difference form; neighbour p formed as its owner does.
    f3 q = {0.0f, 0.0f, 0.0f};
    if (own) {
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      if (reg_ell) {
        // Chunks of 8 neighbours: all their loads are issued before any use.
#pragma unroll
        for (int c0 = 0; c0 < kPcgNbr; c0 += kPcgNbrChunk) {
          if (c0 >= e_w) break;
          f3 zn[kPcgNbrChunk], pp[kPcgNbrChunk];
#pragma unroll
          for (int k = 0; k < kPcgNbrChunk; ++k) {
            zn[k] = ald4(a.zbuf + 4 * static_cast<size_t>(ncol[c0 + k]));
            pp[k] = iter > 0 ? ald4(p_prev + 4 * static_cast<size_t>(ncol[c0 + k]))
                             : f3{0.0f, 0.0f, 0.0f};
          }
#pragma unroll
          for (int k = 0; k < kPcgNbrChunk; ++k) {
            const f3 pn = iter > 0 ? fma3(beta, pp[k], zn[k]) : zn[k];
            const float v = nval[c0 + k];
            ax += v * (pn.x - p.x);
            ay += v * (pn.y - p.y);
            az += v * (pn.z - p.z);
          }
        }
      } else {
        for (int k = 0; k < e_w; ++k) {
          const int c = a.sell_col[e_off + k * kSlice + e_lane];
          const float v = a.sell_val[e_off + k * kSlice + e_lane];
          f3 pn = ald4(a.zbuf + 4 * static_cast<size_t>(c));
          if (iter > 0)
            pn = fma3(beta, ald4(p_prev + 4 * static_cast<size_t>(c)), pn);
          ax += v * (pn.x - p.x);
          ay += v * (pn.y - p.y);
          az += v * (pn.z - p.z);
        }
      }
      q = {ax + dd * p.x, ay + dd * p.y, az + dd * p.z};
    }
    if (coarse) publish_node_sums(q);
    {
      const double s = block_sum(own ? ddot3(p, q) : 0.0, red);
      if (threadIdx.x == 0) astd(my_part + 0, s);
    }
    if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
    const double pAp = wg_total(grp_part, 0, S.nwg, red);
    const float alpha = static_cast<float>(fabs(pAp) < 1e-20 ? 0.0 : rz / pAp);
    dx = fma3(alpha, p, dx);
    r = fma3(-alpha, q, r);
    if (coarse) {
      for (int n = threadIdx.x; n < S.n1; n += kPcgThreads) {
        const f3 qs = ald4(a.nsum + 4 * static_cast<size_t>(S.cn0 + n));
        crl[4 * n] = fmaf(-alpha, qs.x, crl[4 * n]);
        crl[4 * n + 1] = fmaf(-alpha, qs.y, crl[4 * n + 1]);
        crl[4 * n + 2] = fmaf(-alpha, qs.z, crl[4 * n + 2]);
      }
      __syncthreads();
    }
    z = mas(r);
    if (own) ast4(a.zbuf + 4 * static_cast<size_t>(slot), z);
    {
      double s0 = own ? ddot3(r, z) : 0.0, s1 = own ? ddot3(r, r) : 0.0;
      block_sum2(s0, s1, red);
      if (threadIdx.x == 0) {
        astd(my_part + 1, s0);
        astd(my_part + 2, s1);
      }
    }
    if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
    double rz_new;
    wg_total2(grp_part, 1, 2, S.nwg, red, rz_new, rr);
    ++iter;
    if (iter >= a.max_iter) status = 2;
    else if (rr <= abs2) status = 0;
    else if (rr <= rel2 * rr0) status = 1;
    beta = static_cast<float>(fabs(rz) < 1e-20 ? 0.0 : rz_new / rz);
    rz = rz_new;
  }
  if (own) st3(a.x, row, ld3(a.x, row) + dx);
  if (j == 0 && threadIdx.x == 0) {
    PcgState o;
    o.rz = rz;
    o.rr = rr;
    o.rr0 = rr0;
    o.bb = 0.0;
    o.iter = iter;
    o.done = 1;
    o.status = status;
    o.pad = 0;
    a.out[sys] = o;
    if (a.ctl != nullptr) {
      atomicAdd(&a.ctl->pcg_iters, iter);
      if (sys == 0) {
        atomicAdd(&a.ctl->pcg_solves, 1);
        a.ctl->last_rel = static_cast<float>(rr0 > 0.0 ? sqrt(rr / rr0) : 0.0);
      }
    }
  }
}

int g_max_blocks_per_cu = -1;
int g_cus = 0;
size_t g_lds_checked = static_cast<size_t>(-1);

}  // namespace

size_t pcg_persistent_lds(const PersistDev& a) {
  return sizeof(float) * (static_cast<size_t>(kBanksPerWg) * 1024 +
                          static_cast<size_t>(a.ncb_max) * 1024 +
                          4 * static_cast<size_t>(a.n1_max) +
                          8 * static_cast<size_t>(a.ncn_max)) +
         sizeof(int) * (static_cast<size_t>(kBankSlots) * a.ncb_max +
                        static_cast<size_t>(a.ncn_max) + 1 +
                        static_cast<size_t>(a.cnl_max));
}

bool pcg_persistent_fits(int grid, size_t shmem) {
  if (shmem > static_cast<size_t>(kPcgLdsBytes)) return false;
  if (g_max_blocks_per_cu < 0 || g_lds_checked != shmem) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    if (!prop.cooperativeLaunch) return false;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_pcg_persistent),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(shmem)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_pcg_persistent, kPcgThreads, shmem) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    g_max_blocks_per_cu = nb;
    g_cus = prop.multiProcessorCount;
    g_lds_checked = shmem;
  }
  return grid > 0 && grid <= g_max_blocks_per_cu * g_cus;
}

hipError_t launch_pcg_persistent(hipStream_t s, const PersistDev& a,
                                 float rel_tol) {
  PersistDev arg = a;
  float tol = rel_tol;
  void* args[] = {&arg, &tol};
  const size_t shmem = pcg_persistent_lds(a);
  // The grid was sized to fit the device (pcg_persistent_fits) and every
  // barrier wait is bounded, so a plain launch is safe as well: it is used
  // when the cooperative one is refused (the runtime grants cooperative
  // launches to one process of a device at a time) and with SILK_HIP_COOP=0
  // (profilers that do not follow cooperative launches).  Without the
  // runtime's residency guarantee a failure is reported, never a hang.
  static const bool coop_env = [] {
    const char* e = std::getenv("SILK_HIP_COOP");
    return !(e != nullptr && e[0] == '0');
  }();
  if (coop_env) {
    const hipError_t e = hipLaunchCooperativeKernel(
        reinterpret_cast<const void*>(k_pcg_persistent), dim3(a.grid),
        dim3(kPcgThreads), args, static_cast<unsigned>(shmem), s);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
  }
  k_pcg_persistent<<<dim3(a.grid), dim3(kPcgThreads),
                     static_cast<unsigned>(shmem), s>>>(arg, tol);
  return hipGetLastError();
}

}  // namespace silk::amd
