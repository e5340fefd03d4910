// Persistent MAS-preconditioned CG: one launch per solve.
//
// Same iteration as the reference MASCGSolver (mas_cg_solver.cu:193-311:
// alpha and beta guarded at |.| < 1e-20, stop on rr <= abs^2 or
// rr <= rel^2 rr0 or max_iter) and the same additive MAS preconditioner
// (mas_preconditioner.cu:601-728), without a launch per phase.
//
// Work split: a system (cloth) owns a group of workgroups.  A slot workgroup
// owns up to eight consecutive 32-slot level-0 banks, one thread per slot; a
// thread keeps its row's r, dx, z, p, q, neighbour list and row of its bank
// inverse in registers for the whole solve, and the workgroup keeps in LDS p
// of its own slots and of its halo (the other workgroups' slots its rows
// reference).  A coarse workgroup owns no slots: it keeps in LDS the
// inverses of the coarse banks it solves and the level-1 sums they read, and
// solves them while the slot workgroups compute their level-0 part.
//
// Per iteration k:
//   p_k = z_k + beta_k p_{k-1} for the own slots and, with the same
//                     operations, for the halo slots (z0 from the last
//                     exchange plus the coarse corrections of their nodes)
//   q = A p_k         difference form, neighbour p from LDS
//   exchange 1        level-1 node sums of q and the p.q partials (and a
//                     word of each coarse workgroup)
//   alpha; slot workgroups: dx += alpha p; r -= alpha q; z0 = level-0 part
//                     of M r.  Coarse workgroups: the level-1 sums of r
//                     follow by cr -= alpha * qsum; coarse bank cb is solved
//                     by coarse workgroup cb % ncw, which publishes its
//                     members' corrections and the bank's r.z_coarse (the
//                     correction is constant over a node)
//   exchange 2        z0 of the own slots, r.z0 and r.r partials and the
//                     coarse words; then z = z0 + coarse corrections, beta
//                     and the stopping test (every workgroup sums the same
//                     partials in the same order, so all take the same
//                     decision)
//
// An exchange is a set of 64-bit words (epoch << 32 | payload bits) written
// with device-scope relaxed atomic stores; a reader polls the words it needs
// until each carries the exchange's epoch.  A 64-bit atomic is read whole, so
// a word with the right epoch holds the right payload: the words are their
// own barrier, with no fence and no flag.  A word is rewritten only after an
// exchange that each of its readers enters after reading it: exchange-1
// words after the next exchange 2, whose coarse words each coarse workgroup
// publishes after reading exchange 1; exchange-2 words after the next
// exchange 1, whose coarse-workgroup words are published after reading
// exchange 2.
// Epochs continue across launches from a per-system counter.  Waits are
// bounded: a wait that gives up raises an abort flag, every workgroup leaves,
// and the host reports an error.
#include <cstdlib>
#include <utility>

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
    // The flag is read every 256 polls: a read per poll would double the
    // time between two looks at the words.
    if ((++spins & 255) == 0 &&
        (spins > kSpinLimit ||
         __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Partial e (workgroup e / kPcgWaves, wavefront e % kPcgWaves) of the
// block at word offset h of each workgroup's kPcgPartWords (a double whose
// low word comes first; offsets are even and the array 16-byte aligned).
__device__ __forceinline__ double part_at(const float* pw, int e, int h) {
  const int o = kPcgPartWords * (e / kPcgWaves) + h + 2 * (e % kPcgWaves);
  return *reinterpret_cast<const double*>(pw + o);
}

// Sum over the wavefront; the same bits in every lane and in every
// wavefront summing the same values (sum64).
__device__ __forceinline__ double wave_sum(double v) { return sum64(v); }

// Total of the per-wavefront partials of block h over the workgroups plus
// the n_extra doubles at `extra`: lane l adds partials l, l + 64, ... and
// then extras l, l + 64, ... in order, then a butterfly; the same bits in
// every thread of every workgroup.
__device__ __forceinline__ double part_total(const float* pw, int nwg, int h,
                                             const float* extra = nullptr,
                                             int n_extra = 0) {
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  double a = 0.0;
  for (int e = lane; e < kPcgWaves * nwg; e += kWave) a += part_at(pw, e, h);
  for (int e = lane; e < n_extra; e += kWave)
    a += *reinterpret_cast<const double*>(extra + 2 * e);
  return wave_sum(a);
}

// part_total of blocks h0 (with the extras) and h1 at once: the same bits
// as two calls, with the two sums interleaved.
__device__ __forceinline__ void part_total2(const float* pw, int nwg, int h0,
                                            const float* extra, int n_extra,
                                            int h1, double& t0, double& t1) {
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  double a = 0.0, b = 0.0;
  for (int e = lane; e < kPcgWaves * nwg; e += kWave) {
    a += part_at(pw, e, h0);
    b += part_at(pw, e, h1);
  }
  for (int e = lane; e < n_extra; e += kWave)
    a += *reinterpret_cast<const double*>(extra + 2 * e);
  t0 = wave_sum(a);
  t1 = wave_sum(b);
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

// Sum over the 32 lanes of a half-wave (sum32: fixed order).
__device__ __forceinline__ f3 half_sum(f3 v) {
  return {sum32(v.x), sum32(v.y), sum32(v.z)};
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
  __shared__ double red[17];

  int sys = -1;
  for (int s = 0; s < a.nsys; ++s)
    if (static_cast<int>(blockIdx.x) >= a.sys[s].wg0 &&
        static_cast<int>(blockIdx.x) < a.sys[s].wg0 + a.sys[s].nwg + a.sys[s].ncw)
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
  // A slot workgroup (j < nwg) owns level-0 banks; a coarse workgroup
  // (c = j - nwg) owns no slots and solves coarse banks c, c + ncw, ...
  const bool cwg = j >= S.nwg;
  const int c = j - S.nwg;
  const int nb0 = S.nslot / kBankSlots;
  const int b_lo = cwg ? 0 : (j * nb0) / S.nwg;
  const int b_hi = cwg ? 0 : ((j + 1) * nb0) / S.nwg;
  const int lb = b_lo + t / kBankSlots;
  const int hw = t / kBankSlots;
  const int s_in_bank = t & (kBankSlots - 1);
  const bool in_bank = lb < b_hi;
  const int slot = S.slot0 + kBankSlots * lb + s_in_bank;
  const int row = in_bank ? a.slot_row[slot] : -1;
  const bool own = row >= 0;
  const bool coarse = S.ncb > 0;
  const int h0 = a.halo_ptr[blockIdx.x];
  const int nh = a.halo_ptr[blockIdx.x + 1] - h0;
  const size_t hl = static_cast<size_t>(kPcgThreads) + a.nhalo_max;

  float* invc = dyn;                                           // cbw_max x 1024
  float* csum = invc + static_cast<size_t>(a.cbw_max) * 1024;  // 4 x n1_max
  float* czl = csum + 4 * static_cast<size_t>(a.n1_max);      // 4 x ncn_max
  float* qs = czl + 4 * static_cast<size_t>(a.ncn_max);       // 4 x n1_max
  float* cin = qs + 4 * static_cast<size_t>(a.n1_max);        // 4 x 32
  float* cpart = cin + 4 * kBankSlots;                         // 4 x 256
  float* pw = cpart + 4 * kPcgThreads;                         // parts x nwg_max
  float* cpw = pw + kPcgPartWords * static_cast<size_t>(a.nwg_max);  // 4 x ncb_max
  float* zl = cpw + 4 * static_cast<size_t>(a.ncb_max);       // 4 x hl
  float* pl = zl + 4 * hl;                                     // 4 x hl
  int* cbm = reinterpret_cast<int*>(pl + 4 * hl);              // 32 x cbw_max
  int* cnp = cbm + kBankSlots * a.cbw_max;                     // ncn_max + 1
  int* cnl = cnp + a.ncn_max + 1;                              // cnl_max
  int* nl1 = cnl + a.cnl_max;                                  // need1_max
  int* cbn = nl1 + a.need1_max;                                // cbw_max
  const int n1_0 = a.wg_n1_ptr[blockIdx.x];
  const int nn1 = a.wg_n1_ptr[blockIdx.x + 1] - n1_0;
  const int c_0 = a.wg_c_ptr[blockIdx.x];
  const int nc = a.wg_c_ptr[blockIdx.x + 1] - c_0;

  // Static data into LDS: the inverses and members of the coarse banks this
  // workgroup solves (cb = c, c + ncw, ...; none for a slot workgroup),
  // every coarse node's level-1 list, and the level-1 nodes those banks read.
  for (int e = t; cwg && e < a.cbw_max * 1024; e += kPcgThreads) {
    const int cb = c + S.ncw * (e / 1024);
    if (cb < S.ncb)
      invc[e] = a.bank_inv[static_cast<size_t>(a.cb_bank[S.cb0 + cb]) * 1024 + (e & 1023)];
  }
  for (int e = t; e < a.cbw_max * kBankSlots; e += kPcgThreads) {
    const int cb = c + S.ncw * (e / kBankSlots);
    cbm[e] = cwg && cb < S.ncb ? a.cb_member[kBankSlots * (S.cb0 + cb) + e % kBankSlots] : -1;
  }
  const int cn_first = S.ncn > 0 ? a.cn_ptr[S.cn0] : 0;
  for (int e = t; e <= S.ncn; e += kPcgThreads)
    cnp[e] = a.cn_ptr[S.cn0 + e] - cn_first;
  for (int e = t; S.ncn > 0 && e < a.cn_ptr[S.cn0 + S.ncn] - cn_first;
       e += kPcgThreads)
    cnl[e] = a.cn_list[cn_first + e];
  for (int e = t; e < nn1; e += kPcgThreads) nl1[e] = a.wg_n1[n1_0 + e];
  // Members present in each of those banks (they fill the first slots).
  for (int ob = t; ob < a.cbw_max; ob += kPcgThreads) {
    const int cb = c + S.ncw * ob;
    int nm = 0;
    if (cwg && cb < S.ncb)
      while (nm < kBankSlots && a.cb_member[kBankSlots * (S.cb0 + cb) + nm] >= 0) ++nm;
    cbn[ob] = nm;
  }

  // This slot's row of its level-0 bank inverse, starting with the entries of
  // the slots in its own row of 16 lanes: winv[k] is entry k ^ (16 * row).
  float winv[kBankSlots];
  const int row16 = s_in_bank & 16;
#pragma unroll
  for (int k = 0; k < kBankSlots; ++k)
    winv[k] = in_bank ? a.bank_inv[static_cast<size_t>(S.bank0 + lb) * 1024 +
                                   (k ^ row16) * kBankSlots + s_in_bank]
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
  // Coarse nodes of the halo slots this thread completes.
  int hsc[kPcgHaloPerThread][4];
#pragma unroll
  for (int m = 0; m < kPcgHaloPerThread; ++m) {
    const int hh = t + kPcgThreads * m;
#pragma unroll
    for (int l = 0; l < 4; ++l)
      hsc[m][l] = coarse && hh < nh ? a.slot_cn[4 * a.halo_slot[h0 + hh] + l] : -1;
  }
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

  // Exchange 1: the level-1 sums the coarse banks of this workgroup read
  // (3 words per node), the slot workgroups' p.q partials and one word of
  // each coarse workgroup (published once it has read the last exchange 2,
  // so that no word is rewritten before all its readers are done).
  // Exchange 2: the slot workgroups' r.z0 and r.r partials, each coarse
  // bank's r.z_coarse (2 words), and for a slot workgroup z0 of its halo and
  // the corrections of the coarse nodes its slots add (3 words each).
  constexpr int kWaveWords = 2 * kPcgWaves;  // one double per wavefront
  // A coarse workgroup's word: its wavefront 0's (zero) p.q partial.
  constexpr int kAckWord = 0;
  Share x1, x2;
  {
    const int nw1 = 3 * nn1;
    const int aw1 = nw1 + kWaveWords * S.nwg;
    const int w1 = aw1 + S.ncw;
    const int pw2 = 2 * kWaveWords * S.nwg;
    const int cw2 = pw2 + 2 * S.ncb;
    const int hw2 = cw2 + 3 * nh;
    const int w2 = hw2 + 3 * nc;
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
          const int n = a.wg_n1[n1_0 + i / 3];
          x1.addr[m] = a.tn + 3 * static_cast<size_t>(S.cn0 + n) + i % 3;
          x1.dest[m] = static_cast<int>(qs - dyn) + 4 * n + i % 3;
        } else if (i < aw1) {
          const int w = (i - nw1) / kWaveWords, h = (i - nw1) % kWaveWords;
          x1.addr[m] = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + h;
          x1.dest[m] = static_cast<int>(pw - dyn) + kPcgPartWords * w + h;
        } else {
          const int w = S.nwg + (i - aw1);
          x1.addr[m] = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + kAckWord;
          x1.dest[m] = static_cast<int>(pw - dyn) + kPcgPartWords * w + kAckWord;
        }
      }
      x2.addr[m] = a.tp;
      x2.dest[m] = 0;
      if (i < w2) {
        x2.mask |= 1u << m;
        if (i < pw2) {
          const int w = i / (2 * kWaveWords), h = kWaveWords + i % (2 * kWaveWords);
          x2.addr[m] = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + h;
          x2.dest[m] = static_cast<int>(pw - dyn) + kPcgPartWords * w + h;
        } else if (i < cw2) {
          x2.addr[m] = a.tcp + 2 * static_cast<size_t>(S.cb0) + (i - pw2);
          x2.dest[m] = static_cast<int>(cpw - dyn) + (i - pw2);
        } else if (i < hw2) {
          const int hh = (i - cw2) / 3, c = (i - cw2) % 3;
          x2.addr[m] = a.tz + 3 * static_cast<size_t>(a.halo_slot[h0 + hh]) + c;
          x2.dest[m] = static_cast<int>(zl - dyn) + 4 * (kPcgThreads + hh) + c;
        } else {
          const int n = a.wg_c[c_0 + (i - hw2) / 3], c = (i - hw2) % 3;
          x2.addr[m] = a.tc + 3 * static_cast<size_t>(S.cn0 + n) + c;
          x2.dest[m] = static_cast<int>(czl - dyn) + 4 * n + c;
        }
      }
    }
  }
  const int wave = t / kWave;
  const bool lane0 = (t % kWave) == 0;
  unsigned long long* my_tp =
      a.tp + kPcgPartWords * static_cast<size_t>(blockIdx.x) + 2 * wave;
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

  // Level-0 part of z = M r: this slot's bank row times r of its bank (the
  // 32 lanes of its half-wave), r of each slot broadcast from its lane: step
  // m adds slot m of the own row of 16 lanes and slot m of the other row.
  // Called by every thread (all lanes active); lanes outside a bank give 0.
  auto mas_l0 = [&](f3 rv) -> f3 {
    const f3 rs = {xor16_f(rv.x), xor16_f(rv.y), xor16_f(rv.z)};
    f3 z0 = {0.0f, 0.0f, 0.0f};
    [&]<int... M>(std::integer_sequence<int, M...>) {
      ((z0 = {fmaf(winv[M], dpp_f<kDppRowBcast + M>(rv.x), z0.x),
              fmaf(winv[M], dpp_f<kDppRowBcast + M>(rv.y), z0.y),
              fmaf(winv[M], dpp_f<kDppRowBcast + M>(rv.z), z0.z)},
        z0 = {fmaf(winv[16 + M], dpp_f<kDppRowBcast + M>(rs.x), z0.x),
              fmaf(winv[16 + M], dpp_f<kDppRowBcast + M>(rs.y), z0.y),
              fmaf(winv[16 + M], dpp_f<kDppRowBcast + M>(rs.z), z0.z)}),
       ...);
    }(std::make_integer_sequence<int, 16>{});
    return z0;
  };

  // The coarse banks of this workgroup: each member's input is the sum of r
  // over its level-1 nodes (in a level-1 bank, recognised by its first
  // member, which is always present, the level-1 sum itself; above, a group
  // of lanes per member).  The product with the bank inverse is split over the
  // eight half-waves (four inputs each) and joined by the first one, which
  // publishes the members' corrections and the bank's r . z_coarse (the
  // correction is constant over a node).
  auto solve_banks = [&](unsigned e) {
    int ob = 0;
    for (int cb = c; cwg && cb < S.ncb; cb += S.ncw, ++ob) {
      const int* mem = cbm + kBankSlots * ob;
      if (mem[0] < S.n1) {
        // A level-1 bank: its inputs are level-1 sums themselves.
        if (t < kBankSlots) {
          const int m = mem[t];
          sts4(cin, t, m >= 0 ? lds4(csum, m) : f3{0.0f, 0.0f, 0.0f});
        }
      } else {
        // Members are the first nm slots; each gets g lanes (8 to 64, as
        // many as the bank's members leave), the level-1 nodes strided.
        const int nm = cbn[ob];
        const int g = nm <= 4 ? 64 : (nm <= 8 ? 32 : (nm <= 16 ? 16 : 8));
        const int k = t / g, sub = t % g;
        const int m = k < nm ? mem[k] : -1;
        f3 v = {0.0f, 0.0f, 0.0f};
        if (m >= 0) {
          const int e_end = cnp[m + 1];
          for (int ee = cnp[m] + sub; ee < e_end; ee += 4 * g) {
            int id[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) id[u] = ee + g * u < e_end ? cnl[ee + g * u] : -1;
            f3 c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
              c[u] = id[u] >= 0 ? lds4(csum, id[u]) : f3{0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (id[u] >= 0) v = v + c[u];
          }
        }
        // Sum over the lane group (the same g in the whole workgroup).
        v = {sum8(v.x), sum8(v.y), sum8(v.z)};
        if (g >= 16) {
          v.x += dpp_f<kDppMirror>(v.x);
          v.y += dpp_f<kDppMirror>(v.y);
          v.z += dpp_f<kDppMirror>(v.z);
        }
        if (g >= 32) v = {v.x + xor16_f(v.x), v.y + xor16_f(v.y), v.z + xor16_f(v.z)};
        if (g >= 64) v = {v.x + xor32_f(v.x), v.y + xor32_f(v.y), v.z + xor32_f(v.z)};
        if (sub == 0 && k < nm) sts4(cin, k, v);
        if (t >= nm && t < kBankSlots) sts4(cin, t, f3{0.0f, 0.0f, 0.0f});
      }
      __syncthreads();
      {
        const float* inv = invc + static_cast<size_t>(ob) * 1024;
        const int k0 = 4 * hw;
        f3 part = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = k0; k < k0 + 4; ++k) {
          const float w = inv[k * kBankSlots + s_in_bank];
          const f3 c = lds4(cin, k);
          part.x += w * c.x;
          part.y += w * c.y;
          part.z += w * c.z;
        }
        sts4(cpart, t, part);
      }
      __syncthreads();
      if (t < kBankSlots) {
        f3 zc = lds4(cpart, t);
#pragma unroll
        for (int q = 1; q < kBanksPerWg; ++q) zc = zc + lds4(cpart, q * kBankSlots + t);
        const int mi = mem[t];
        if (mi >= 0) put_f3(a.tc + 3 * static_cast<size_t>(S.cn0 + mi), e, zc);
        const double d = sum32(mi >= 0 ? ddot3(lds4(cin, t), zc) : 0.0);
        if (t == 0) put_double(a.tcp + 2 * static_cast<size_t>(S.cb0 + cb), e, d);
      }
      if (cb + S.ncw < S.ncb) __syncthreads();
    }
  };

  // z of an own slot: the level-0 part plus the coarse nodes' corrections,
  // in level order; halo slots are completed with the same additions.
  auto add_coarse = [&](f3 z0, const int* cn) -> f3 {
#pragma unroll
    for (int l = 0; l < 4; ++l)
      if (cn[l] >= 0) z0 = z0 + lds4(czl, cn[l]);
    return z0;
  };

  // Exchange 2: a slot workgroup computes z0 = level-0 part of M r and
  // publishes z0 of its slots and its wavefronts' r.z0 and r.r; a coarse
  // workgroup solves its banks.  Then every workgroup collects the exchange;
  // sets z (slot workgroups) and the totals.  False when the exchange gave
  // up.
  auto finish_z = [&](f3& z_out, double& rz_out, double& rr_out) -> bool {
    ++ep;
    f3 z0 = {0.0f, 0.0f, 0.0f};
    if (cwg) {
      solve_banks(ep);
    } else {
      z0 = mas_l0(r);
      if (own) put_f3(a.tz + 3 * static_cast<size_t>(slot), ep, z0);
      const double s0 = wave_sum(own ? ddot3(r, z0) : 0.0);
      const double s1 = wave_sum(own ? ddot3(r, r) : 0.0);
      if (lane0) {
        put_double(my_tp + kWaveWords, ep, s0);
        put_double(my_tp + 2 * kWaveWords, ep, s1);
      }
    }
    if (!__syncthreads_and(collect(x2, ep, dyn, a.abort_flag))) return false;
    z_out = coarse && in_bank ? add_coarse(z0, scn) : z0;
    // r.z = r.z0 over the slots plus r.z_coarse over the coarse banks.
    part_total2(pw, S.nwg, kWaveWords, cpw, S.ncb, 2 * kWaveWords, rz_out, rr_out);
    return true;
  };

  // ---- prologue: level-1 sums of r0, z0 = M r0, r.z, r.r
  if (coarse) {
    ++ep;
    publish_node_sums(r, ep);
    if (lane0) put_double(my_tp, ep, 0.0);
    if (!__syncthreads_and(collect(x1, ep, dyn, a.abort_flag))) return;
    for (int e = t; e < nn1; e += kPcgThreads) sts4(csum, nl1[e], lds4(qs, nl1[e]));
  }
  double rz, rr;
  f3 z;
  if (!finish_z(z, rz, rr)) return;
  const double rr0 = rr;
  const double abs2 = static_cast<double>(a.abs_tol) * a.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  int status = isfinite(rz) ? -1 : 3;
  int iter = 0;
  float beta = 0.0f;

  while (status < 0) {
    p = (iter == 0) ? z : fma3(beta, p, z);
    sts4(pl, t, p);
#pragma unroll
    for (int m = 0; m < kPcgHaloPerThread; ++m) {
      const int hh = t + kPcgThreads * m;
      if (hh < nh) {
        const f3 zh = coarse ? add_coarse(lds4(zl, kPcgThreads + hh), hsc[m])
                             : lds4(zl, kPcgThreads + hh);
        sts4(pl, kPcgThreads + hh,
             iter == 0 ? zh : fma3(beta, lds4(pl, kPcgThreads + hh), zh));
      }
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
      const double s = wave_sum(own ? ddot3(p, q) : 0.0);
      if (lane0) put_double(my_tp, ep, s);
    }
    if (!__syncthreads_and(collect(x1, ep, dyn, a.abort_flag))) return;
    const double pAp = part_total(pw, S.nwg, 0);
    const float alpha = static_cast<float>(fabs(pAp) < 1e-20 ? 0.0 : rz / pAp);
    dx = fma3(alpha, p, dx);
    r = fma3(-alpha, q, r);
    for (int e = t; e < nn1; e += kPcgThreads) {
      const int n = nl1[e];
      const f3 qn = lds4(qs, n), cn = lds4(csum, n);
      sts4(csum, n, {fmaf(-alpha, qn.x, cn.x), fmaf(-alpha, qn.y, cn.y),
                     fmaf(-alpha, qn.z, cn.z)});
    }
    double rz_new;
    if (!finish_z(z, rz_new, rr)) return;
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
  return sizeof(float) * (static_cast<size_t>(a.cbw_max) * 1024 +
                          8 * static_cast<size_t>(a.n1_max) +
                          4 * static_cast<size_t>(a.ncn_max) + 4 * kBankSlots +
                          4 * kPcgThreads +
                          kPcgPartWords * static_cast<size_t>(a.nwg_max) +
                          4 * static_cast<size_t>(a.ncb_max) + 8 * hl) +
         sizeof(int) * (static_cast<size_t>(kBankSlots) * a.cbw_max +
                        static_cast<size_t>(a.ncn_max) + 1 +
                        static_cast<size_t>(a.cnl_max) +
                        static_cast<size_t>(a.need1_max) +
                        static_cast<size_t>(a.cbw_max));
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
