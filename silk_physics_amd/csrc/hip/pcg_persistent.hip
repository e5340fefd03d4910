// Persistent MAS-preconditioned CG: one cooperative launch per solve.
//
// Same iteration as the reference MASCGSolver (mas_cg_solver.cu:193-311:
// alpha and beta guarded at |.| < 1e-20, stop on rr <= abs^2 or
// rr <= rel^2 rr0 or max_iter) and the same additive MAS preconditioner
// (mas_preconditioner.cu:601-728), but without a launch per phase: the rows
// stay in registers of the workgroup that owns their bank for the whole
// solve and workgroups synchronise through grid barriers.
//
// Per iteration k (p_k = z_k + beta_k p_{k-1}):
//   q = A p_k        neighbour values p_k(j) are formed on the fly from z_k(j)
//                    and p_{k-1}(j), both published before the previous
//                    barrier, so p needs no barrier of its own
//   B1  p.q          -> alpha; dx += alpha p; r -= alpha q
//   level-0 banks solved from the workgroup's own rows (LDS exchange);
//   level-1 sums of r published
//   B2  coarse banks solved: redundantly by every workgroup into LDS when
//       they are few, otherwise shared out and followed by one more barrier
//   z = z0 + coarse z of the row's nodes; z published; r.z and r.r
//   B3  beta and the stopping test: every workgroup sums the same partials
//       in the same order, so all take the same decision
#include "hip/device_common.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

constexpr unsigned kSpinLimit = 1u << 22;

// Grid barrier of one workgroup group; count and gen sit in different
// 128-byte lines.  Returns false when the wait gave up (a workgroup never
// arrived, or another one gave up and raised abort_flag): every caller then
// leaves the kernel, so every wave ends.
__device__ bool group_barrier(unsigned* count, unsigned* gen, unsigned nwg,
                              unsigned& epoch, int* abort_flag, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned arrived = __hip_atomic_fetch_add(
        count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    epoch += 1u;
    if (!ok)
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

// Sum of a[stride * i], i < n, in a fixed order; valid in every thread.
__device__ double wg_total(const double* a, int stride, int n, double* lds) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += a[stride * i];
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

__global__ void __launch_bounds__(kPcgThreads)
    k_pcg_persistent(PersistDev a, float rel_tol) {
  extern __shared__ float cz_lds[];  // 3 * lds_nodes when coarse_lds
  __shared__ float rb[kPcgThreads][3];
  __shared__ float cbuf[kPcgThreads][3];
  __shared__ int cn1[kPcgThreads];
  __shared__ double red[17];
  __shared__ int s_ok;

  int sys = -1;
  for (int s = 0; s < a.nsys; ++s)
    if (static_cast<int>(blockIdx.x) >= a.sys[s].wg0 &&
        static_cast<int>(blockIdx.x) < a.sys[s].wg0 + a.sys[s].nwg)
      sys = s;
  if (sys < 0) return;  // spare workgroup: takes no part in any barrier
  // Stopped ADMM loop: nothing to solve.  ctl is not written before every
  // workgroup has passed its first barrier, so all read the same value.
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
  const int s_in_bank = threadIdx.x & (kBankSlots - 1);
  const int base = threadIdx.x & ~(kBankSlots - 1);
  const bool in_bank = lb < b_hi;
  const int slot = S.slot0 + kBankSlots * lb + s_in_bank;
  const int row = in_bank ? a.slot_row[slot] : -1;
  const bool own = row >= 0;
  const bool coarse = S.ncb > 0;
  const size_t ns3 = 3 * static_cast<size_t>(a.nslot_total);

  int e_off = 0, e_w = 0, e_lane = 0;
  if (own) {
    const int sl = slot / kSlice;
    e_off = a.sell_off[sl];
    e_w = a.sell_width[sl];
    e_lane = slot % kSlice;
  }
  const float dd = own ? a.d[row] : 0.0f;
  f3 r = own ? ld3(a.r0, row) : f3{0.0f, 0.0f, 0.0f};
  f3 dx = {0.0f, 0.0f, 0.0f};
  f3 p = {0.0f, 0.0f, 0.0f};
  cn1[threadIdx.x] = (in_bank && coarse) ? a.slot_cn[4 * slot] : -1;

  double* my_part = a.part + 4 * blockIdx.x;
  const double* grp_part = a.part + 4 * S.wg0;

  // z = M r for the rows of this workgroup (takes 1 or 2 barriers when the
  // system has coarse levels).
  auto mas = [&](f3 rv, bool& ok) -> f3 {
    rb[threadIdx.x][0] = rv.x;
    rb[threadIdx.x][1] = rv.y;
    rb[threadIdx.x][2] = rv.z;
    __syncthreads();
    f3 z0 = {0.0f, 0.0f, 0.0f};
    if (in_bank) {
      const float* inv = a.bank_inv + static_cast<size_t>(S.bank0 + lb) * 1024;
      for (int k = 0; k < kBankSlots; ++k) {
        const float w = inv[k * kBankSlots + s_in_bank];
        z0.x += w * rb[base + k][0];
        z0.y += w * rb[base + k][1];
        z0.z += w * rb[base + k][2];
      }
    }
    if (!coarse) return z0;
    // Level-1 sums: the first slot of each level-1 node of the bank sums it.
    const int n = cn1[threadIdx.x];
    if (n >= 0) {
      bool leader = true;
      for (int k = 0; k < s_in_bank; ++k) leader = leader && cn1[base + k] != n;
      if (leader) {
        f3 s = {0.0f, 0.0f, 0.0f};
        for (int k = s_in_bank; k < kBankSlots; ++k)
          if (cn1[base + k] == n)
            s = s + f3{rb[base + k][0], rb[base + k][1], rb[base + k][2]};
        st3(a.cr, S.cn0 + n, s);
      }
    }
    ok = group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok);
    if (!ok) return z0;
    // Coarse banks: gather the 32 member sums, multiply by the inverse.
    const int hw = threadIdx.x / kBankSlots;
    const int per_round = kPcgThreads / kBankSlots;
    const int first = a.coarse_lds ? 0 : j * per_round;
    const int stride = a.coarse_lds ? per_round : S.nwg * per_round;
    for (int round = first; round < S.ncb; round += stride) {
      const int cb = round + hw;
      int m = -1;
      f3 v = {0.0f, 0.0f, 0.0f};
      if (cb < S.ncb) {
        m = a.cb_member[kBankSlots * (S.cb0 + cb) + s_in_bank];
        if (m >= 0) {
          const int g = S.cn0 + m;
          for (int e = a.cn_ptr[g]; e < a.cn_ptr[g + 1]; ++e)
            v = v + ld3(a.cr, S.cn0 + a.cn_list[e]);
        }
      }
      cbuf[threadIdx.x][0] = v.x;
      cbuf[threadIdx.x][1] = v.y;
      cbuf[threadIdx.x][2] = v.z;
      __syncthreads();
      if (cb < S.ncb && m >= 0) {
        const float* inv =
            a.bank_inv + static_cast<size_t>(a.cb_bank[S.cb0 + cb]) * 1024;
        f3 zc = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < kBankSlots; ++k) {
          const float w = inv[k * kBankSlots + s_in_bank];
          zc.x += w * cbuf[base + k][0];
          zc.y += w * cbuf[base + k][1];
          zc.z += w * cbuf[base + k][2];
        }
        if (a.coarse_lds) st3(cz_lds, m, zc);
        else st3(a.cz, S.cn0 + m, zc);
      }
      __syncthreads();
    }
    if (!a.coarse_lds) {
      ok = group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok);
      if (!ok) return z0;
    }
    if (in_bank) {
      for (int l = 0; l < 4; ++l) {
        const int c = a.slot_cn[4 * slot + l];
        if (c < 0) break;
        z0 = z0 + (a.coarse_lds ? ld3(cz_lds, c) : ld3(a.cz, S.cn0 + c));
      }
    }
    return z0;
  };

  bool ok = true;
  f3 z = mas(r, ok);
  if (!ok) return;
  if (own) st3(a.zbuf, slot, z);
  {
    const double s0 = block_sum(own ? ddot3(r, z) : 0.0, red);
    const double s1 = block_sum(own ? ddot3(r, r) : 0.0, red);
    if (threadIdx.x == 0) {
      my_part[1] = s0;
      my_part[2] = s1;
    }
  }
  if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
  double rz = wg_total(grp_part + 1, 4, S.nwg, red);
  double rr = wg_total(grp_part + 2, 4, S.nwg, red);
  const double rr0 = rr;
  const double abs2 = static_cast<double>(a.abs_tol) * a.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  int status = isfinite(rz) ? -1 : 3;
  int iter = 0;
  float beta = 0.0f;

  while (status < 0) {
    const float* p_prev = a.pbuf + static_cast<size_t>((iter + 1) & 1) * ns3;
    float* p_cur = a.pbuf + static_cast<size_t>(iter & 1) * ns3;
    p = (iter == 0) ? z : fma3(beta, p, z);
    if (own) st3(p_cur, slot, p);
    // q = (K + D) p in difference form; neighbour p formed as its owner does.
    f3 q = {0.0f, 0.0f, 0.0f};
    if (own) {
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      for (int k = 0; k < e_w; ++k) {
        const int c = a.sell_col[e_off + k * kSlice + e_lane];
        const float v = a.sell_val[e_off + k * kSlice + e_lane];
        f3 pn = ld3(a.zbuf, c);
        if (iter > 0) pn = fma3(beta, ld3(p_prev, c), pn);
        ax += v * (pn.x - p.x);
        ay += v * (pn.y - p.y);
        az += v * (pn.z - p.z);
      }
      q = {ax + dd * p.x, ay + dd * p.y, az + dd * p.z};
    }
    {
      const double s = block_sum(own ? ddot3(p, q) : 0.0, red);
      if (threadIdx.x == 0) my_part[0] = s;
    }
    if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
    const double pAp = wg_total(grp_part + 0, 4, S.nwg, red);
    const float alpha = static_cast<float>(fabs(pAp) < 1e-20 ? 0.0 : rz / pAp);
    dx = fma3(alpha, p, dx);
    r = fma3(-alpha, q, r);
    z = mas(r, ok);
    if (!ok) return;
    if (own) st3(a.zbuf, slot, z);
    {
      const double s0 = block_sum(own ? ddot3(r, z) : 0.0, red);
      const double s1 = block_sum(own ? ddot3(r, r) : 0.0, red);
      if (threadIdx.x == 0) {
        my_part[1] = s0;
        my_part[2] = s1;
      }
    }
    if (!group_barrier(bcount, bgen, W, epoch, a.abort_flag, &s_ok)) return;
    const double rz_new = wg_total(grp_part + 1, 4, S.nwg, red);
    rr = wg_total(grp_part + 2, 4, S.nwg, red);
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

bool pcg_persistent_fits(int grid, int lds_nodes) {
  const size_t shmem = 3 * sizeof(float) * static_cast<size_t>(lds_nodes);
  if (g_max_blocks_per_cu < 0 || g_lds_checked != shmem) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    if (!prop.cooperativeLaunch) return false;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_pcg_persistent, kPcgThreads, shmem) != hipSuccess)
      return false;
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
  const size_t shmem =
      a.coarse_lds ? 3 * sizeof(float) * static_cast<size_t>(a.lds_nodes) : 0;
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pcg_persistent),
                                    dim3(a.grid), dim3(kPcgThreads), args,
                                    static_cast<unsigned>(shmem), s);
}

}  // namespace silk::amd
