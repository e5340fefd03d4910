// Persistent MAS-preconditioned CG: one launch per solve.
//
// Same iteration as the reference MASCGSolver (mas_cg_solver.cu:193-311:
// alpha and beta guarded at |.| < 1e-20, stop on rr <= abs^2 or
// rr <= rel^2 rr0 or max_iter) and the same additive MAS preconditioner
// (mas_preconditioner.cu:601-728), without a launch per phase.
//
// Work split: a system (cloth) owns a group of workgroups; a workgroup owns
// up to eight consecutive 32-slot level-0 banks, one thread per slot.  A
// thread keeps its row's r, dx, z, p, q, neighbour list and row of its bank
// inverse in registers for the whole solve.  The workgroup keeps in LDS every
// coarse bank inverse of the system, the coarse sums, and p of its own slots
// and of its halo (the other workgroups' slots its rows reference).
//
// Per iteration k:
//   p_k = z_k + beta_k p_{k-1} for the own slots and, with the same fmaf,
//                     for the halo slots (their z came with the last exchange)
//   q = A p_k         difference form, neighbour p from LDS
//   exchange 1        level-1 node sums of q and the p.q partials
//   alpha; dx += alpha p; r -= alpha q; the level-1 sums of r follow by
//                     cr -= alpha * qsum; z = M r: level-0 banks from the
//                     workgroup's rows, coarse levels from the level-1 sums
//   exchange 2        z of the own slots, r.z and r.r partials; beta and the
//                     stopping test (every workgroup sums the same partials
//                     in the same order, so all take the same decision)
//
// An exchange is a set of 64-bit words (epoch << 32 | payload bits) written
// with device-scope relaxed atomic stores; a reader polls the words it needs
// until each carries the exchange's epoch.  A 64-bit atomic is read whole, so
// a word with the right epoch holds the right payload: the words are their
// own barrier, with no fence and no flag.  A word is rewritten only after an
// exchange that each of its readers enters after reading it (node sums and
// p.q after the next exchange 2; z, r.z and r.r after the next exchange 1).
// Epochs continue across launches from a per-system counter.  Waits are
// bounded: a wait that gives up raises an abort flag, every workgroup leaves,
// and the host reports an error.
#include <cstdlib>

#include "hip/device_common.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

constexpr unsigned kSpinLimit = 1u << 22;
constexpr int kBanksPerWg = kPcgThreads / kBankSlots;  // 8

__device__ __forceinline__ void put_word(unsigned long long* w, unsigned epoch,
                                         unsigned bits) {
  __hip_atomic_store(w, (static_cast<unsigned long long>(epoch) << 32) | bits,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void put_f3(unsigned long long* w, unsigned epoch,
                                       f3 v) {
  put_word(w, epoch, __float_as_uint(v.x));
  put_word(w + 1, epoch, __float_as_uint(v.y));
  put_word(w + 2, epoch, __float_as_uint(v.z));
}

__device__ __forceinline__ void put_double(unsigned long long* w, unsigned epoch,
                                           double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  put_word(w, epoch, static_cast<unsigned>(b));
  put_word(w + 1, epoch, static_cast<unsigned>(b >> 32));
}

// A thread's share of an exchange: word i of the workgroup's list belongs to
// thread i % kPcgThreads; addr is the tagged word, dest the float of the LDS
// array that receives its payload.
struct Share {
  const unsigned long long* addr[kPcgXchWords];
  int dest[kPcgXchWords];
  unsigned mask;
};

// Waits until each of the thread's words carries `epoch`; stores the
// payloads.  Returns false when the wait gave up.
__device__ bool collect(const Share& sh, unsigned epoch, float* lds,
                        int* abort_flag) {
  unsigned pending = sh.mask;
  unsigned spins = 0;
  while (pending != 0) {
    unsigned long long v[kPcgXchWords];
#pragma unroll
    for (int m = 0; m < kPcgXchWords; ++m) {
      v[m] = 0;
      if (pending & (1u << m))
        v[m] = __hip_atomic_load(const_cast<unsigned long long*>(sh.addr[m]),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int m = 0; m < kPcgXchWords; ++m)
      if ((pending & (1u << m)) && static_cast<unsigned>(v[m] >> 32) == epoch) {
        lds[sh.dest[m]] = __uint_as_float(static_cast<unsigned>(v[m]));
        pending &= ~(1u << m);
      }
    if (pending == 0) break;
    if (++spins > kSpinLimit ||
        __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__device__ __forceinline__ double part_at(const float* pw, int w, int h) {
  const unsigned long long lo = __float_as_uint(pw[8 * w + h]);
  const unsigned long long hi = __float_as_uint(pw[8 * w + h + 1]);
  return __longlong_as_double(static_cast<long long>((hi << 32) | lo));
}

// Sums over the workgroups of the doubles received as word pairs at
// pw[8 w + h], pw[8 w + h + 1] (h = h0, h1): lane l adds workgroups l,
// l + 64, ... in order, then a butterfly over the wavefront.  Each step adds
// the same two partial sums in every lane (in either order), so every lane
// of every wavefront of every workgroup gets the same bits.
__device__ __forceinline__ void part_totals(const float* pw, int nwg, int h0,
                                            int h1, double& t0, double& t1) {
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  double a = 0.0, b = 0.0;
  for (int w = lane; w < nwg; w += kWave) {
    a += part_at(pw, w, h0);
    b += part_at(pw, w, h1);
  }
  for (int m = kWave / 2; m > 0; m >>= 1) {
    a += __shfl_xor(a, m, kWave);
    b += __shfl_xor(b, m, kWave);
  }
  t0 = a;
  t1 = b;
}

__device__ __forceinline__ f3 fma3(float a, f3 x, f3 y) {
  return {fmaf(a, x.x, y.x), fmaf(a, x.y, y.y), fmaf(a, x.z, y.z)};
}

__device__ __forceinline__ double ddot3(f3 a, f3 b) {
  return static_cast<double>(a.x) * b.x + static_cast<double>(a.y) * b.y +
         static_cast<double>(a.z) * b.z;
}

// x, y, z of the 16-byte record i (one LDS read).
__device__ __forceinline__ f3 lds4(const float* base, int i) {
  const float4 v = *reinterpret_cast<const float4*>(base + 4 * i);
  return {v.x, v.y, v.z};
}

__device__ __forceinline__ void sts4(float* base, int i, f3 v) {
  base[4 * i] = v.x;
  base[4 * i + 1] = v.y;
  base[4 * i + 2] = v.z;
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

__global__ void __launch_bounds__(kPcgThreads)
    k_pcg_persistent(PersistDev a, float rel_tol) {
  // Dynamic LDS (floats): [ncb_max coarse inverses][coarse sums: level-1
  // nodes first][coarse z][q sums of the level-1 nodes][partials, 8 per
  // workgroup][z of own + halo slots][p of own + halo slots] (4 floats per
  // node or slot), then the coarse bank members, the descendant-list
  // offsets and the lists (ints).
  extern __shared__ __align__(16) float dyn[];
  __shared__ __align__(16) float rb[kPcgThreads][4];
  __shared__ double red[17];

  int sys = -1;
  for (int s = 0; s < a.nsys; ++s)
    if (static_cast<int>(blockIdx.x) >= a.sys[s].wg0 &&
        static_cast<int>(blockIdx.x) < a.sys[s].wg0 + a.sys[s].nwg)
      sys = s;
  if (sys < 0) return;  // spare workgroup: takes no part in any exchange
  // Stopped ADMM loop: nothing to solve.  ctl->done is written only by
  // other kernels, so every workgroup reads the same value.
  if (a.ctl != nullptr && a.ctl->done != 0) return;
  if (a.ctl != nullptr) rel_tol = a.ctl->rel_tol;
  const PcgSysDev S = a.sys[sys];
  // The previous solve's last epoch; workgroup 0 advances it at the end,
  // after every workgroup has read it (all take part in the first exchange).
  unsigned* epoch_ctr = a.bar + kPcgBarrierStride * sys;
  unsigned ep = __hip_atomic_load(epoch_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int t = static_cast<int>(threadIdx.x);
  const int j = static_cast<int>(blockIdx.x) - S.wg0;
  const int nb0 = S.nslot / kBankSlots;
  const int b_lo = (j * nb0) / S.nwg, b_hi = ((j + 1) * nb0) / S.nwg;
  const int lb = b_lo + t / kBankSlots;
  const int hw = t / kBankSlots;
  const int s_in_bank = t & (kBankSlots - 1);
  const int base = t & ~(kBankSlots - 1);
  const bool in_bank = lb < b_hi;
  const int slot = S.slot0 + kBankSlots * lb + s_in_bank;
  const int row = in_bank ? a.slot_row[slot] : -1;
  const bool own = row >= 0;
  const bool coarse = S.ncb > 0;
  const int h0 = a.halo_ptr[blockIdx.x];
  const int nh = a.halo_ptr[blockIdx.x + 1] - h0;
  const size_t hl = static_cast<size_t>(kPcgThreads) + a.nhalo_max;

  float* invc = dyn;                                           // ncb_max x 1024
  float* csum = invc + static_cast<size_t>(a.ncb_max) * 1024;  // 4 x ncn_max
  float* czl = csum + 4 * static_cast<size_t>(a.ncn_max);     // 4 x ncn_max
  float* qs = czl + 4 * static_cast<size_t>(a.ncn_max);       // 4 x n1_max
  float* pw = qs + 4 * static_cast<size_t>(a.n1_max);         // 8 x nwg_max
  float* zl = pw + 8 * static_cast<size_t>(a.nwg_max);        // 4 x hl
  float* pl = zl + 4 * hl;                                     // 4 x hl
  int* cbm = reinterpret_cast<int*>(pl + 4 * hl);              // 32 x ncb_max
  int* cnp = cbm + kBankSlots * a.ncb_max;                     // ncn_max + 1
  int* cnl = cnp + a.ncn_max + 1;                              // cnl_max

  // Static data into LDS.
  for (int e = t; e < S.ncb * 1024; e += kPcgThreads)
    invc[e] = a.bank_inv[static_cast<size_t>(a.cb_bank[S.cb0 + e / 1024]) * 1024 +
                         (e & 1023)];
  for (int e = t; e < S.ncb * kBankSlots; e += kPcgThreads)
    cbm[e] = a.cb_member[kBankSlots * S.cb0 + e];
  const int cn_first = S.ncn > 0 ? a.cn_ptr[S.cn0] : 0;
  for (int e = t; e <= S.ncn; e += kPcgThreads)
    cnp[e] = a.cn_ptr[S.cn0 + e] - cn_first;
  for (int e = t; S.ncn > 0 && e < a.cn_ptr[S.cn0 + S.ncn] - cn_first;
       e += kPcgThreads)
    cnl[e] = a.cn_list[cn_first + e];

  // This slot's row of its level-0 bank inverse.
  float winv[kBankSlots];
#pragma unroll
  for (int k = 0; k < kBankSlots; ++k)
    winv[k] = in_bank ? a.bank_inv[static_cast<size_t>(S.bank0 + lb) * 1024 +
                                   k * kBankSlots + s_in_bank]
                      : 0.0f;
  int e_off = 0, e_w = 0, e_lane = 0;
  if (own) {
    const int sl = slot / kSlice;
    e_off = a.sell_off[sl];
    e_w = a.sell_width[sl];
    e_lane = slot % kSlice;
  }
  // This row's neighbours as indices into pl (slot-ELL, padded to whole
  // chunks with the row's own slot and weight 0) and coarse nodes, kept in
  // registers.
  const bool reg_ell = e_w <= kPcgNbr;
  int ncol[kPcgNbr];
  float nval[kPcgNbr];
#pragma unroll
  for (int k = 0; k < kPcgNbr; ++k) {
    const bool v = own && reg_ell && k < e_w;
    ncol[k] = v ? a.sell_lcol[e_off + k * kSlice + e_lane] : t;
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

  // Exchange 1: level-1 sums (3 words per node) then p.q (2 per workgroup).
  // Exchange 2: r.z, r.r (4 per workgroup) then z of the halo (3 per slot).
  Share x1, x2;
  {
    const int nw1 = coarse ? 3 * S.n1 : 0;
    const int w1 = nw1 + 2 * S.nwg;
    const int pw2 = 4 * S.nwg;
    const int w2 = pw2 + 3 * nh;
    x1.mask = 0;
    x2.mask = 0;
#pragma unroll
    for (int m = 0; m < kPcgXchWords; ++m) {
      const int i = t + kPcgThreads * m;
      x1.addr[m] = a.tp;
      x1.dest[m] = 0;
      if (i < w1) {
        x1.mask |= 1u << m;
        if (i < nw1) {
          x1.addr[m] = a.tn + 3 * static_cast<size_t>(S.cn0 + i / 3) + i % 3;
          x1.dest[m] = static_cast<int>(qs - dyn) + 4 * (i / 3) + i % 3;
        } else {
          const int w = (i - nw1) / 2, h = (i - nw1) % 2;
          x1.addr[m] = a.tp + 8 * static_cast<size_t>(S.wg0 + w) + h;
          x1.dest[m] = static_cast<int>(pw - dyn) + 8 * w + h;
        }
      }
      x2.addr[m] = a.tp;
      x2.dest[m] = 0;
      if (i < w2) {
        x2.mask |= 1u << m;
        if (i < pw2) {
          const int w = i / 4, h = i % 4;
          x2.addr[m] = a.tp + 8 * static_cast<size_t>(S.wg0 + w) + 2 + h;
          x2.dest[m] = static_cast<int>(pw - dyn) + 8 * w + 2 + h;
        } else {
          const int hh = (i - pw2) / 3, c = (i - pw2) % 3;
          x2.addr[m] = a.tz + 3 * static_cast<size_t>(a.halo_slot[h0 + hh]) + c;
          x2.dest[m] = static_cast<int>(zl - dyn) + 4 * (kPcgThreads + hh) + c;
        }
      }
    }
  }
  unsigned long long* my_tp = a.tp + 8 * static_cast<size_t>(blockIdx.x);
  __syncthreads();

  // Sum of v over the slots of this thread's level-1 node, published by the
  // first slot of the node (level-1 nodes never straddle banks).
  auto publish_node_sums = [&](f3 v, unsigned e) {
    for (int k = 0; k < n1_count; ++k) {
      const int target = n1_first + k;
      const f3 s = half_sum(scn[0] == target ? v : f3{0.0f, 0.0f, 0.0f});
      if (s_in_bank == 0) put_f3(a.tn + 3 * static_cast<size_t>(S.cn0 + target), e, s);
    }
  };

  // z = M r from this thread's r, its bank row and the level-1 sums in csum.
  auto mas = [&](f3 rv) -> f3 {
    rb[t][0] = rv.x;
    rb[t][1] = rv.y;
    rb[t][2] = rv.z;
    __syncthreads();
    f3 z0 = {0.0f, 0.0f, 0.0f};
    if (in_bank) {
#pragma unroll
      for (int k = 0; k < kBankSlots; ++k) {
        const float w = winv[k];
        const f3 rk = lds4(&rb[0][0], base + k);
        z0.x += w * rk.x;
        z0.y += w * rk.y;
        z0.z += w * rk.z;
      }
    }
    if (!coarse) return z0;
    // Sums of the deeper coarse nodes from their level-1 nodes: one
    // half-wave per node, the level-1 nodes strided over its lanes.
    for (int n = S.n1 + hw; n < S.ncn; n += kBanksPerWg) {
      const int e_end = cnp[n + 1];
      f3 v = {0.0f, 0.0f, 0.0f};
      for (int e = cnp[n] + s_in_bank; e < e_end; e += 4 * kBankSlots) {
        int id[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          id[u] = e + u * kBankSlots < e_end ? cnl[e + u * kBankSlots] : -1;
        f3 c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          c[u] = id[u] >= 0 ? lds4(csum, id[u]) : f3{0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (id[u] >= 0) v = v + c[u];
      }
      v = half_sum(v);
      if (s_in_bank == 0) sts4(csum, n, v);
    }
    __syncthreads();
    // Coarse banks: one thread per (bank, member); the bank's inputs are
    // read by the whole half-wave at once.
    // All loads of a dot product are issued before its first product.
    for (int wk = t; wk < S.ncb * kBankSlots; wk += kPcgThreads) {
      const int cb = wk / kBankSlots, i = wk % kBankSlots;
      const float* inv = invc + static_cast<size_t>(cb) * 1024;
      const int* mem = cbm + kBankSlots * cb;
      f3 zc = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k0 = 0; k0 < kBankSlots; k0 += 16) {
        int mk[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) mk[k] = mem[k0 + k];
        f3 v[16];
        float w[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          v[k] = mk[k] >= 0 ? lds4(csum, mk[k]) : f3{0.0f, 0.0f, 0.0f};
          w[k] = inv[(k0 + k) * kBankSlots + i];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          zc.x += w[k] * v[k].x;
          zc.y += w[k] * v[k].y;
          zc.z += w[k] * v[k].z;
        }
      }
      if (mem[i] >= 0) sts4(czl, mem[i], zc);
    }
    __syncthreads();
    if (in_bank) {
#pragma unroll
      for (int l = 0; l < 4; ++l)
        if (scn[l] >= 0) z0 = z0 + lds4(czl, scn[l]);
    }
    return z0;
  };

  // ---- prologue: level-1 sums of r0, z0 = M r0, r.z, r.r
  if (coarse) {
    ++ep;
    publish_node_sums(r, ep);
    if (t == 0) put_double(my_tp, ep, 0.0);
    if (!__syncthreads_and(collect(x1, ep, dyn, a.abort_flag))) return;
    for (int n = t; n < S.n1; n += kPcgThreads) sts4(csum, n, lds4(qs, n));
  }
  f3 z = mas(r);
  ++ep;
  if (own) put_f3(a.tz + 3 * static_cast<size_t>(slot), ep, z);
  {
    double s0 = own ? ddot3(r, z) : 0.0, s1 = own ? ddot3(r, r) : 0.0;
    block_sum2(s0, s1, red);
    if (t == 0) {
      put_double(my_tp + 2, ep, s0);
      put_double(my_tp + 4, ep, s1);
    }
  }
  if (!__syncthreads_and(collect(x2, ep, dyn, a.abort_flag))) return;
  double rz, rr;
  part_totals(pw, S.nwg, 2, 4, rz, rr);
  const double rr0 = rr;
  const double abs2 = static_cast<double>(a.abs_tol) * a.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  int status = isfinite(rz) ? -1 : 3;
  int iter = 0;
  float beta = 0.0f;

  while (status < 0) {
    p = (iter == 0) ? z : fma3(beta, p, z);
    sts4(pl, t, p);
    for (int hh = t; hh < nh; hh += kPcgThreads) {
      const f3 zh = lds4(zl, kPcgThreads + hh);
      sts4(pl, kPcgThreads + hh,
           iter == 0 ? zh : fma3(beta, lds4(pl, kPcgThreads + hh), zh));
    }
    __syncthreads();
    // q = (K + D) p in difference form.
    f3 q = {0.0f, 0.0f, 0.0f};
    if (own) {
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      if (reg_ell) {
#pragma unroll
        for (int c0 = 0; c0 < kPcgNbr; c0 += kPcgNbrChunk) {
          if (c0 >= e_w) break;
          f3 pn[kPcgNbrChunk];
#pragma unroll
          for (int k = 0; k < kPcgNbrChunk; ++k) pn[k] = lds4(pl, ncol[c0 + k]);
#pragma unroll
          for (int k = 0; k < kPcgNbrChunk; ++k) {
            const float v = nval[c0 + k];
            ax += v * (pn[k].x - p.x);
            ay += v * (pn[k].y - p.y);
            az += v * (pn[k].z - p.z);
          }
        }
      } else {
        for (int k = 0; k < e_w; ++k) {
          const f3 pn = lds4(pl, a.sell_lcol[e_off + k * kSlice + e_lane]);
          const float v = a.sell_val[e_off + k * kSlice + e_lane];
          ax += v * (pn.x - p.x);
          ay += v * (pn.y - p.y);
          az += v * (pn.z - p.z);
        }
      }
      q = {ax + dd * p.x, ay + dd * p.y, az + dd * p.z};
    }
    ++ep;
    if (coarse) publish_node_sums(q, ep);
    {
      const double s = block_sum(own ? ddot3(p, q) : 0.0, red);
      if (t == 0) put_double(my_tp, ep, s);
    }
    if (!__syncthreads_and(collect(x1, ep, dyn, a.abort_flag))) return;
    double pAp, unused;
    part_totals(pw, S.nwg, 0, 0, pAp, unused);
    const float alpha = static_cast<float>(fabs(pAp) < 1e-20 ? 0.0 : rz / pAp);
    dx = fma3(alpha, p, dx);
    r = fma3(-alpha, q, r);
    if (coarse)
      for (int n = t; n < S.n1; n += kPcgThreads) {
        const f3 qn = lds4(qs, n), cn = lds4(csum, n);
        sts4(csum, n, {fmaf(-alpha, qn.x, cn.x), fmaf(-alpha, qn.y, cn.y),
                       fmaf(-alpha, qn.z, cn.z)});
      }
    z = mas(r);
    ++ep;
    if (own) put_f3(a.tz + 3 * static_cast<size_t>(slot), ep, z);
    {
      double s0 = own ? ddot3(r, z) : 0.0, s1 = own ? ddot3(r, r) : 0.0;
      block_sum2(s0, s1, red);
      if (t == 0) {
        put_double(my_tp + 2, ep, s0);
        put_double(my_tp + 4, ep, s1);
      }
    }
    if (!__syncthreads_and(collect(x2, ep, dyn, a.abort_flag))) return;
    double rz_new;
    part_totals(pw, S.nwg, 2, 4, rz_new, rr);
    ++iter;
    if (iter >= a.max_iter) status = 2;
    else if (rr <= abs2) status = 0;
    else if (rr <= rel2 * rr0) status = 1;
    beta = static_cast<float>(fabs(rz) < 1e-20 ? 0.0 : rz_new / rz);
    rz = rz_new;
  }
  if (own) st3(a.x, row, ld3(a.x, row) + dx);
  if (j == 0 && t == 0) {
    __hip_atomic_store(epoch_ctr, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  const size_t hl = static_cast<size_t>(kPcgThreads) + a.nhalo_max;
  return sizeof(float) * (static_cast<size_t>(a.ncb_max) * 1024 +
                          8 * static_cast<size_t>(a.ncn_max) +
                          4 * static_cast<size_t>(a.n1_max) +
                          8 * static_cast<size_t>(a.nwg_max) + 8 * hl) +
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
  // wait is bounded, so a plain launch is safe as well: it is used when the
  // cooperative one is refused (the runtime grants cooperative launches to
  // one process of a device at a time) and with SILK_HIP_COOP=0 (profilers
  // that do not follow cooperative launches).  Without the runtime's
  // residency guarantee a failure is reported, never a hang.
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
