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
// reference).  A coarse workgroup owns no slots: each of its wavefronts
// solves one coarse bank, keeping its members' residual sums (in double) and
// its rows of the bank inverse in registers, while the slot workgroups
// compute their level-0 part.  Bank products broadcast the inputs across
// lanes with DPP (row_newbcast) instead of staging them in LDS.
//
// Per iteration k:
//   p_k = z_k + beta_k p_{k-1} for the own slots and, with the same
//                     operations, for the halo slots (z0 from the last
//                     exchange plus the coarse corrections of their nodes)
//   q = A p_k         difference form, neighbour p from LDS
//   exchange 1        level-1 node sums of q and the p.q partials (and a
//                     word of each coarse workgroup)
//   alpha; slot workgroups: dx += alpha p; r -= alpha q; z0 = level-0 part
//                     of M r.  Coarse workgroups: a member's residual sum
//                     follows by cr -= alpha * (its sum of the level-1 q
//                     sums); coarse bank cb is solved by wavefront cb / ncw
//                     of coarse workgroup cb % ncw, which publishes its
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

#include "hip/device_common.hpp"
#include "hip/dpp_dot.hpp"
#include "hip/kernels.hpp"

namespace silk::amd {

using namespace dev;

namespace {

constexpr unsigned kSpinLimit = 1u << 22;

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

// Words of an exchange mapped when collected: thread t polls words
// first, first + kPcgThreads, ... < total, kPcgXchBatch at a time, mapping
// each with word(i, &addr, &dest) when its batch starts.  Used for the words
// beyond the kPcgThreads * kPcgXchWords kept in registers (a coarse bank
// that reads the level-1 sums of a whole system, a large halo; first =
// t + kPcgThreads * kPcgXchWords) and for the exchanges of the true-residual
// test (first = t).  Returns false when the wait gave up.
template <class WordFn>
__device__ bool collect_rest(int first, int total, WordFn&& word, unsigned epoch,
                             float* lds, int* abort_flag) {
  constexpr int kStride = kPcgThreads * kPcgXchBatch;
  for (int i0 = first; i0 < total; i0 += kStride) {
    const unsigned long long* addr[kPcgXchBatch];
    int dest[kPcgXchBatch];
    unsigned pending = 0;
#pragma unroll
    for (int m = 0; m < kPcgXchBatch; ++m) {
      const int i = i0 + kPcgThreads * m;
      addr[m] = nullptr;
      dest[m] = 0;
      if (i < total) {
        word(i, addr[m], dest[m]);
        pending |= 1u << m;
      }
    }
    unsigned spins = 0;
    while (pending != 0) {
      unsigned long long v[kPcgXchBatch];
#pragma unroll
      for (int m = 0; m < kPcgXchBatch; ++m) {
        v[m] = 0;
        if (pending & (1u << m))
          v[m] = __hip_atomic_load(const_cast<unsigned long long*>(addr[m]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int m = 0; m < kPcgXchBatch; ++m)
        if ((pending & (1u << m)) && static_cast<unsigned>(v[m] >> 32) == epoch) {
          lds[dest[m]] = __uint_as_float(static_cast<unsigned>(v[m]));
          pending &= ~(1u << m);
        }
      if (pending == 0) break;
      if ((++spins & 255) == 0 &&
          (spins > kSpinLimit ||
           __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return true;
}

// Sum over the wavefront; the same bits in every lane and in every
// wavefront summing the same values (sum64).
__device__ __forceinline__ double wave_sum(double v) { return sum64(v); }

// Adds to `a` the doubles src[e] (two floats each in LDS, low word first,
// 8-byte aligned) for e = lane, lane + 64, ... < n in that order; the loads
// of four steps are issued before their additions.
__device__ __forceinline__ void add_strided(double& a, const float* src, int n) {
  const double* d = reinterpret_cast<const double*>(src);
  const int lane = static_cast<int>(threadIdx.x % kWave);
  for (int e0 = lane; e0 < n; e0 += 4 * kWave) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = d[min(e0 + u * kWave, n - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e0 + u * kWave < n) a += v[u];
  }
}

// Total of n per-wavefront partials (partial e of wavefront e % kPcgWaves
// of workgroup e / kPcgWaves) plus the n_extra doubles at `extra`: lane l
// adds partials l, l + 64, ... and then extras l, l + 64, ... in order,
// then a butterfly; the same bits in every thread of every workgroup.
__device__ __forceinline__ double part_total(const float* parts, int n,
                                             const float* extra = nullptr,
                                             int n_extra = 0) {
  double a = 0.0;
  if (n <= 2 * kWave && n_extra <= kWave) {
    // Up to two partials and one extra per lane: all loads at once.
    const double* d = reinterpret_cast<const double*>(parts);
    const double* x = reinterpret_cast<const double*>(extra);
    const int lane = static_cast<int>(threadIdx.x % kWave);
    const double v0 = d[min(lane, n - 1)], v1 = d[min(lane + kWave, n - 1)];
    const double v2 = n_extra > 0 ? x[min(lane, n_extra - 1)] : 0.0;
    if (lane < n) a += v0;
    if (lane + kWave < n) a += v1;
    if (lane < n_extra) a += v2;
  } else {
    add_strided(a, parts, n);
    add_strided(a, extra, n_extra);
  }
  return wave_sum(a);
}

// part_total of two sets of partials (the first with the extras): the same
// bits as two calls, with the two sums interleaved.
__device__ __forceinline__ void part_total2(const float* pa, const float* pb, int n,
                                            const float* extra, int n_extra,
                                            double& ta, double& tb) {
  double a = 0.0, b = 0.0;
  if (n <= 2 * kWave && n_extra <= kWave) {
    // Up to two partials of each set and one extra per lane: all loads at
    // once (the additions in the order of the general case).
    const double* da = reinterpret_cast<const double*>(pa);
    const double* db = reinterpret_cast<const double*>(pb);
    const double* x = reinterpret_cast<const double*>(extra);
    const int lane = static_cast<int>(threadIdx.x % kWave);
    const int i0 = min(lane, n - 1), i1 = min(lane + kWave, n - 1);
    const double a0 = da[i0], a1 = da[i1], b0 = db[i0], b1 = db[i1];
    const double x0 = n_extra > 0 ? x[min(lane, n_extra - 1)] : 0.0;
    if (lane < n) {
      a += a0;
      b += b0;
    }
    if (lane + kWave < n) {
      a += a1;
      b += b1;
    }
    if (lane < n_extra) a += x0;
  } else {
    add_strided(a, pa, n);
    add_strided(b, pb, n);
    add_strided(a, extra, n_extra);
  }
  sum64_pair(a, b);
  ta = a;
  tb = b;
}

__device__ __forceinline__ f3 fma3(float a, f3 x, f3 y) {
  return {fmaf(a, x.x, y.x), fmaf(a, x.y, y.y), fmaf(a, x.z, y.z)};
}

// Three doubles: a coarse member's residual sum.  A sum over thousands of
// slots kept as a float would carry a rounding of its initial size, which
// near convergence exceeds the sum itself.
struct d3 {
  double x, y, z;
};

__device__ __forceinline__ double ddot3(f3 a, f3 b) {
  return static_cast<double>(a.x) * b.x + static_cast<double>(a.y) * b.y +
         static_cast<double>(a.z) * b.z;
}

// x, y, z of the 16-byte record i (one LDS read).
__device__ __forceinline__ f3 lds4(const float* base, int i) {
  const float4 v = *reinterpret_cast<const float4*>(base + 4 * i);
  return {v.x, v.y, v.z};
}

// Records of the level-1 sums: 16 bytes (one LDS read) in the
// one-workgroup-per-CU variant, 12 bytes in the two-per-CU variant, whose
// workgroups must fit half of the CU's LDS.
template <int kWavesPerSimd>
constexpr int kQsStride = kWavesPerSimd == 2 ? 3 : 4;

template <int kStride>
__device__ __forceinline__ f3 ldsq(const float* base, int i) {
  if constexpr (kStride == 4) {
    return lds4(base, i);
  } else {
    return {base[kStride * i], base[kStride * i + 1], base[kStride * i + 2]};
  }
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

// kWavesPerSimd 1: one workgroup per CU (no register cap below the
// kernel's ~270 VGPRs); 2: the compiler keeps the kernel within 256 VGPRs
// (a few spills) so that two workgroups share a CU, for systems whose grid
// exceeds the CU count.
template <int kWavesPerSimd>
__global__ void __launch_bounds__(kPcgThreads, kWavesPerSimd)
    k_pcg_persistent(PersistDev a, float rel_tol) {
  // Dynamic LDS (floats): [ncb_max coarse inverses][coarse sums: level-1
  // nodes first][coarse z][q sums of the level-1 nodes][partials, 8 per
  // workgroup][z of own + halo slots][p of own + halo slots] (4 floats per
  // node or slot), then the coarse bank members, the descendant-list
  // offsets and the lists (ints).
  extern __shared__ __align__(16) float dyn[];
  __shared__ double red[17];

  // Workgroup id over all systems: a launch covers the systems of one batch
  // (workgroups [wg_base, wg_base + grid)).
  const int bx = static_cast<int>(blockIdx.x) + a.wg_base;
  int sys = -1;
  for (int s = 0; s < a.nsys; ++s)
    if (bx >= a.sys[s].wg0 && bx < a.sys[s].wg0 + a.sys[s].nwg + a.sys[s].ncw) sys = s;
  if (sys < 0) return;  // spare workgroup: takes no part in any exchange
  // Stopped ADMM loop: nothing to solve.  ctl->done is written only by
  // other kernels, so every workgroup reads the same value.
  if (a.ctl != nullptr && a.ctl->done != 0) return;
  // A restart leaves converged systems alone (sys_done is written by the
  // verification kernel before this launch, so all workgroups agree).
  if (a.sys_done != nullptr && a.sys_done[sys] != 0) return;
  if (a.ctl != nullptr) rel_tol = a.ctl->rel_tol;
  const PcgSysDev S = a.sys[sys];
  // The previous solve's last epoch; workgroup 0 advances it at the end,
  // after every workgroup has read it (all take part in the first exchange).
  unsigned* epoch_ctr = a.bar + kPcgBarrierStride * sys;
  unsigned ep = __hip_atomic_load(epoch_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int t = static_cast<int>(threadIdx.x);
  const int j = bx - S.wg0;
  // A slot workgroup (j < nwg) owns level-0 banks; a coarse workgroup
  // (c = j - nwg) owns no slots and solves coarse banks c, c + ncw, ...
  const bool cwg = j >= S.nwg;
  const int c = j - S.nwg;
  const int nb0 = S.nslot / kBankSlots;
  const int b_lo = cwg ? 0 : (j * nb0) / S.nwg;
  const int b_hi = cwg ? 0 : ((j + 1) * nb0) / S.nwg;
  const int lb = b_lo + t / kBankSlots;
  const int s_in_bank = t & (kBankSlots - 1);
  const bool in_bank = lb < b_hi;
  const int slot = S.slot0 + kBankSlots * lb + s_in_bank;
  const int row = in_bank ? a.slot_row[slot] : -1;
  const bool own = row >= 0;
  const bool coarse = S.ncb > 0;
  const int h0 = a.halo_ptr[bx];
  const int nh = a.halo_ptr[bx + 1] - h0;
  const size_t hl = static_cast<size_t>(kPcgThreads) + a.nhalo_max;

  float* czl = dyn;                                            // 4 x (nc_max + 1)
  float* qs = czl + 4 * static_cast<size_t>(a.nc_max + 1);    // 4 x need_max
  // Per-wavefront partials as doubles (two floats, partial e = wavefront
  // e % kPcgWaves of workgroup e / kPcgWaves): p.q, r.z0 and r.r.
  constexpr int kQs = kQsStride<kWavesPerSimd>;
  float* ppq = qs + ((kQs * static_cast<size_t>(a.need_max) + 3) & ~size_t{3});  // 8 x nwg_max
  float* prz = ppq + 8 * static_cast<size_t>(a.nwg_max);      // 8 x nwg_max
  float* prr = prz + 8 * static_cast<size_t>(a.nwg_max);      // 8 x nwg_max
  float* cpw = prr + 8 * static_cast<size_t>(a.nwg_max);      // 4 x ncb_max
  float* zl = cpw + 4 * static_cast<size_t>(a.ncb_max);       // 4 x hl
  float* pl = zl + 4 * hl;                                     // 4 x hl
  const int n1_0 = a.wg_n1_ptr[bx];
  const int nn1 = a.wg_n1_ptr[bx + 1] - n1_0;
  const int c_0 = a.wg_c_ptr[bx];
  const int nc = a.wg_c_ptr[bx + 1] - c_0;
  const int wave = t / kWave;
  const int lane = t % kWave;
  const bool lane0 = lane == 0;

  // Coarse workgroup: wavefront w solves coarse bank cbk = c + ncw * w (at
  // most one per wavefront).  Lane l keeps the residual sum of member l % 32
  // (cs) and computes output member i_out = l % 16 + 16 * (l / 32) over the
  // inputs kh + [0, 16), kh = l & 16; lanes l and l ^ 16 join their halves.
  const int cbk = c + S.ncw * wave;
  const bool has_bank = cwg && cbk < S.ncb;
  const int i_out = (lane & 15) + 16 * (lane >> 5);
  const int kh = lane & 16;
  const int cbase = kBankSlots * (S.cb0 + (has_bank ? cbk : 0));
  const int mem_s = has_bank ? a.cb_member[cbase + (lane & 31)] : -1;
  const int mem_o = has_bank ? a.cb_member[cbase + i_out] : -1;
  // Positions [rb, re) of the level-1 sums of member l % 32 in this
  // workgroup's list (qs), the members present (they fill the first slots)
  // and whether some member reads more than one sum (a bank above level 1).
  const int rb = has_bank ? a.cb_rng[2 * (cbase + (lane & 31))] : 0;
  const int re = has_bank ? a.cb_rng[2 * (cbase + (lane & 31)) + 1] : 0;
  const int nm = __popcll(__ballot(lane < kBankSlots && mem_s >= 0));
  const bool deep = __ballot(re - rb > 1) != 0;
  if (t < 4) czl[4 * a.nc_max + t] = 0.0f;

  // This slot's row of its level-0 bank inverse, starting with the entries of
  // the slots in its own row of 16 lanes: winv[k] is entry k ^ (16 * row).
  // In a coarse workgroup, winv[m] (m < 16) is entry kh + m of row i_out of
  // the wavefront's coarse bank inverse.
  float winv[kBankSlots];
  const int row16 = s_in_bank & 16;
  // Entry e of bank bk's inverse (16-bit storage when mas_half is set).
  auto inv_at = [&](int bk, size_t e) -> float {
    const size_t i = static_cast<size_t>(bk) * 1024 + e;
    return a.mas_half != 0
               ? a.bank_scale[bk] * half_bits_to_float(a.bank_inv_h[i], a.mas_half)
               : a.bank_inv[i];
  };
  const int cbank = has_bank ? a.cb_bank[S.cb0 + cbk] : 0;
#pragma unroll
  for (int k = 0; k < kBankSlots; ++k)
    winv[k] = in_bank ? inv_at(S.bank0 + lb, (k ^ row16) * kBankSlots + s_in_bank)
              : has_bank && k < 16 ? inv_at(cbank, (kh + k) * kBankSlots + i_out)
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
  // Coarse nodes of the slot (and of the halo slots this thread completes)
  // per level.  czl holds the corrections of the workgroup's own list of
  // coarse nodes (wg_c), so the additions index it by list position; an
  // absent level points at record nc_max, which stays zero, so that the
  // additions need no branch.
  const int zrec = a.nc_max;
  int scn[4] = {-1, -1, -1, -1};
  int scz[4] = {zrec, zrec, zrec, zrec};
  if (in_bank && coarse)
    for (int l = 0; l < 4; ++l) {
      scn[l] = a.slot_cn[4 * slot + l];
      const int q = a.slot_cnl[4 * slot + l];
      scz[l] = q >= 0 ? q : zrec;
    }
  int hsc[kPcgHaloPerThread][4];
#pragma unroll
  for (int m = 0; m < kPcgHaloPerThread; ++m) {
    const int hh = t + kPcgThreads * m;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int q = coarse && hh < nh ? a.halo_cnl[4 * (h0 + hh) + l] : -1;
      hsc[m][l] = q >= 0 ? q : zrec;
    }
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
  const int nw1 = 3 * nn1;
  const int aw1 = nw1 + kWaveWords * S.nwg;
  const int w1 = aw1 + S.ncw;
  const int pw2 = 2 * kWaveWords * S.nwg;
  const int cw2 = pw2 + 2 * S.ncb;
  const int hw2 = cw2 + 3 * nh;
  const int w2 = hw2 + 3 * nc;
  // Word i of each exchange: its tagged word and the float of dyn that
  // receives its payload.
  auto x1_word = [&](int i, const unsigned long long*& addr, int& dest) {
    if (i < nw1) {
      const int n = a.wg_n1[n1_0 + i / 3];
      addr = a.tn + 3 * static_cast<size_t>(n) + i % 3;
      dest = static_cast<int>(qs - dyn) + kQs * (i / 3) + i % 3;
    } else if (i < aw1) {
      const int w = (i - nw1) / kWaveWords, h = (i - nw1) % kWaveWords;
      addr = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + h;
      dest = static_cast<int>(ppq - dyn) + kWaveWords * w + h;
    } else {
      const int w = S.nwg + (i - aw1);
      addr = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + kAckWord;
      dest = static_cast<int>(ppq - dyn) + kWaveWords * w + kAckWord;
    }
  };
  auto x2_word = [&](int i, const unsigned long long*& addr, int& dest) {
    if (i < pw2) {
      const int w = i / (2 * kWaveWords), h = i % (2 * kWaveWords);
      addr = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + kWaveWords + h;
      dest = static_cast<int>((h < kWaveWords ? prz : prr) - dyn) + kWaveWords * w +
             h % kWaveWords;
    } else if (i < cw2) {
      addr = a.tcp + 2 * static_cast<size_t>(S.cb0) + (i - pw2);
      dest = static_cast<int>(cpw - dyn) + (i - pw2);
    } else if (i < hw2) {
      const int hh = (i - cw2) / 3, c = (i - cw2) % 3;
      addr = a.tz + 3 * static_cast<size_t>(a.halo_slot[h0 + hh]) + c;
      dest = static_cast<int>(zl - dyn) + 4 * (kPcgThreads + hh) + c;
    } else {
      const int j = (i - hw2) / 3, c = (i - hw2) % 3;
      addr = a.tc + 3 * static_cast<size_t>(S.cn0 + a.wg_c[c_0 + j]) + c;
      dest = static_cast<int>(czl - dyn) + 4 * j + c;
    }
  };
  Share x1, x2;
  x1.mask = 0;
  x2.mask = 0;
#pragma unroll
  for (int m = 0; m < kPcgXchWords; ++m) {
    const int i = t + kPcgThreads * m;
    x1.addr[m] = a.tp;
    x1.dest[m] = 0;
    if (i < w1) {
      x1.mask |= 1u << m;
      x1_word(i, x1.addr[m], x1.dest[m]);
    }
    x2.addr[m] = a.tp;
    x2.dest[m] = 0;
    if (i < w2) {
      x2.mask |= 1u << m;
      x2_word(i, x2.addr[m], x2.dest[m]);
    }
  }
  // Both exchanges, the words held in registers and the rest.
  const int rest0 = t + kPcgThreads * kPcgXchWords;
  auto collect1 = [&](unsigned e) {
    return collect(x1, e, dyn, a.abort_flag) &&
           collect_rest(rest0, w1, x1_word, e, dyn, a.abort_flag);
  };
  auto collect2 = [&](unsigned e) {
    return collect(x2, e, dyn, a.abort_flag) &&
           collect_rest(rest0, w2, x2_word, e, dyn, a.abort_flag);
  };
  unsigned long long* my_tp =
      a.tp + kPcgPartWords * static_cast<size_t>(bx) + 2 * wave;
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
  // 32 lanes of its half-wave), r of each slot broadcast from its lane: the
  // slots of the own row of 16 lanes, then those of the other row.  Called
  // by every thread (all lanes active); lanes outside a bank give 0.
  auto mas_l0 = [&](f3 rv) -> f3 {
    const f3 rs = {xor16_f(rv.x), xor16_f(rv.y), xor16_f(rv.z)};
    return {row_bcast_dot16(row_bcast_dot16(0.0f, rv.x, winv), rs.x, winv + 16),
            row_bcast_dot16(row_bcast_dot16(0.0f, rv.y, winv), rs.y, winv + 16),
            row_bcast_dot16(row_bcast_dot16(0.0f, rv.z, winv), rs.z, winv + 16)};
  };

  // Coarse bank of this wavefront (called by all its lanes, wave-uniformly).
  // sum_nodes(src): for lane l, the sum (in double) of src over the level-1
  // nodes of member l % 32, positions [rb, re) of the list (zero for an
  // absent member): in a level-1 bank the member's own entry; above, g lanes
  // per member (as many as the members leave) sum its range strided, with
  // the loads of four steps issued together, and lane l takes its member's
  // group sum.
  const int grp = nm <= 1 ? 64 : nm <= 2 ? 32 : nm <= 4 ? 16 : nm <= 8 ? 8 : nm <= 16 ? 4 : 2;
  auto sum_nodes = [&](const float* src) -> d3 {
    if (!deep) {
      const f3 v = rb < re ? ldsq<kQs>(src, rb) : f3{0.0f, 0.0f, 0.0f};
      return {v.x, v.y, v.z};
    }
    const int k = lane / grp, sub = lane % grp;
    const int kb = __shfl(rb, k & 31), ke0 = __shfl(re, k & 31);
    const int ke = k < nm ? ke0 : kb;
    d3 v = {0.0, 0.0, 0.0};
    for (int e0 = kb + sub; e0 < ke; e0 += 4 * grp) {
      f3 cv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) cv[u] = ldsq<kQs>(src, min(e0 + grp * u, ke - 1));
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e0 + grp * u < ke) v = {v.x + cv[u].x, v.y + cv[u].y, v.z + cv[u].z};
    }
    // Sum over the lane group (grp is the same in the whole wavefront).
    auto step = [&](auto&& other) { v = {v.x + other(v.x), v.y + other(v.y), v.z + other(v.z)}; };
    if (grp >= 2) step([](double x) { return dpp_d<kDppXor1>(x); });
    if (grp >= 4) step([](double x) { return dpp_d<kDppXor2>(x); });
    if (grp >= 8) step([](double x) { return dpp_d<kDppHalfMirror>(x); });
    if (grp >= 16) step([](double x) { return dpp_d<kDppMirror>(x); });
    if (grp >= 32) step([](double x) { return xor16_d(x); });
    if (grp >= 64) step([](double x) { return xor32_d(x); });
    const int from = ((lane & 31) * grp) & (kWave - 1);
    const d3 own_sum = {__shfl(v.x, from), __shfl(v.y, from), __shfl(v.z, from)};
    return rb < re ? own_sum : d3{0.0, 0.0, 0.0};
  };

  // zc = bank inverse times the members' residual sums cs (cs of member
  // kh + m broadcast from lane 16 * (l / 16) + m), halves joined over l ^ 16
  // (both lanes then hold the same bits); publishes the members'
  // corrections (the correction is constant over a node) and the bank's
  // r . z_coarse.
  auto solve_bank = [&](d3 csd, unsigned e) {
    const f3 cs = {static_cast<float>(csd.x), static_cast<float>(csd.y),
                   static_cast<float>(csd.z)};
    f3 zc = {row_bcast_dot16(0.0f, cs.x, winv), row_bcast_dot16(0.0f, cs.y, winv),
             row_bcast_dot16(0.0f, cs.z, winv)};
    zc = {zc.x + xor16_f(zc.x), zc.y + xor16_f(zc.y), zc.z + xor16_f(zc.z)};
    if (kh == 0 && mem_o >= 0) put_f3(a.tc + 3 * static_cast<size_t>(S.cn0 + mem_o), e, zc);
    // Lane l contributes member i_out when it also holds that member's sum
    // (lanes 0-15 and 48-63); with at most 16 members only lanes 0-15 do,
    // so the sum spans as many lanes as the members need.
    double d = (lane & 31) == i_out && mem_o >= 0 ? ddot3(cs, zc) : 0.0;
    if (nm > 16) {
      d = sum64(d);
    } else if (nm > 1) {
      d = sum8(d);
      if (nm > 8) d += dpp_d<kDppMirror>(d);
    }
    if (lane0) put_double(a.tcp + 2 * static_cast<size_t>(S.cb0 + cbk), e, d);
  };

  // z of an own slot: the level-0 part plus the coarse nodes' corrections,
  // in level order; halo slots are completed with the same additions.
  auto add_coarse = [&](f3 z0, const int* cn) -> f3 {
    f3 cv[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) cv[l] = lds4(czl, cn[l]);
#pragma unroll
    for (int l = 0; l < 4; ++l) z0 = z0 + cv[l];
    return z0;
  };

  // Residual sum of this lane's coarse member (coarse workgroups).
  d3 cs = {0.0, 0.0, 0.0};

  // Exchange 2: a slot workgroup computes z0 = level-0 part of M r and
  // publishes z0 of its slots and its wavefronts' r.z0 and r.r; a coarse
  // workgroup solves its banks.  Then every workgroup collects the exchange;
  // sets z (slot workgroups) and the totals.  False when the exchange gave
  // up.
  auto finish_z = [&](f3& z_out, double& rz_out, double& rr_out) -> bool {
    ++ep;
    f3 z0 = {0.0f, 0.0f, 0.0f};
    if (cwg) {
      if (has_bank) solve_bank(cs, ep);
    } else {
      z0 = mas_l0(r);
      if (own) put_f3(a.tz + 3 * static_cast<size_t>(slot), ep, z0);
      double s0 = own ? ddot3(r, z0) : 0.0;
      double s1 = own ? ddot3(r, r) : 0.0;
      sum64_pair(s0, s1);
      if (lane0) {
        put_double(my_tp + kWaveWords, ep, s0);
        put_double(my_tp + 2 * kWaveWords, ep, s1);
      }
    }
    if (!__syncthreads_and(collect2(ep))) return false;
    z_out = coarse && in_bank ? add_coarse(z0, scz) : z0;
    // r.z = r.z0 over the slots plus r.z_coarse over the coarse banks.
    part_total2(prz, prr, kPcgWaves * S.nwg, cpw, S.ncb, rz_out, rr_out);
    return true;
  };

  // ---- start, and restart after a residual replacement: level-1 sums of
  // r, z = M r, r.z and r.r
  const double abs2 = static_cast<double>(a.abs_tol) * a.abs_tol;
  const double rel2 = static_cast<double>(rel_tol) * rel_tol;
  double rz = 0.0, rr = 0.0, rr0 = 0.0, ref2 = 0.0;
  f3 z = {0.0f, 0.0f, 0.0f};
  int status = -1, iter = 0;
  float beta = 0.0f;
  // start: the search (re)starts from r; fresh: p = z (the first iteration
  // after a start).
  bool start = true, fresh = true;
  int replaced = 0, unverified = 0;
  double true_ratio = -1.0;

  for (;;) {
    if (start) {
      if (coarse) {
        ++ep;
        publish_node_sums(r, ep);
        if (lane0) put_double(my_tp, ep, 0.0);
        if (!__syncthreads_and(collect1(ep))) return;
        if (has_bank) cs = sum_nodes(qs);
      }
      if (!finish_z(z, rz, rr)) return;
      if (iter == 0 && replaced == 0) {
        rr0 = rr;
        // Reference of the relative test: rr0, or |b|^2 summed from the
        // residual kernel's block partials (every wavefront adds the same
        // values in the same order, so all take the same decisions).
        ref2 = rr0;
        if (a.criterion >= 1) {
          const int q0 = a.sys_blk0[sys], q1 = a.sys_blk0[sys + 1];
          double v = 0.0;
          for (int q = q0 + lane; q < q1; q += kWave) v += a.part2[2 * q + 1];
          ref2 = wave_sum(v);
        }
      }
      status = isfinite(rz) ? -1 : 3;
      // Right-hand-side criterion: a start that already meets it takes no
      // iteration.
      if (status < 0 && a.criterion >= 1) {
        if (rr <= abs2) status = 0;
        else if (rr <= rel2 * ref2) status = 1;
      }
      start = false;
      fresh = true;
    }
    if (status < 0) {
      p = fresh ? z : fma3(beta, p, z);
      sts4(pl, t, p);
#pragma unroll
      for (int m = 0; m < kPcgHaloPerThread; ++m) {
        const int hh = t + kPcgThreads * m;
        if (hh < nh) {
          const f3 zh = coarse ? add_coarse(lds4(zl, kPcgThreads + hh), hsc[m])
                               : lds4(zl, kPcgThreads + hh);
          sts4(pl, kPcgThreads + hh,
               fresh ? zh : fma3(beta, lds4(pl, kPcgThreads + hh), zh));
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
      if (!__syncthreads_and(collect1(ep))) return;
      // The members' sums of q (independent of alpha); their residual sums
      // follow by cs -= alpha * qsum.
      d3 cq = {0.0, 0.0, 0.0};
      if (has_bank) cq = sum_nodes(qs);
      const double pAp = part_total(ppq, kPcgWaves * S.nwg);
      const float alpha = static_cast<float>(fabs(pAp) < 1e-20 ? 0.0 : rz / pAp);
      dx = fma3(alpha, p, dx);
      r = fma3(-alpha, q, r);
      {
        const double am = -static_cast<double>(alpha);
        cs = {fma(am, cq.x, cs.x), fma(am, cq.y, cs.y), fma(am, cq.z, cs.z)};
      }
      double rz_new;
      if (!finish_z(z, rz_new, rr)) return;
      ++iter;
      fresh = false;
      if (iter >= a.max_iter) status = 2;
      else if (rr <= abs2) status = 0;
      else if (rr <= rel2 * ref2) status = 1;
      beta = static_cast<float>(fabs(rz) < 1e-20 ? 0.0 : rz_new / rz);
      rz = rz_new;
      continue;
    }
    // True-residual test of the iterate x0 + dx (x0 is not read: r0d holds
    // b - A x0 in FP64).  Exchange 3: dx of the own slots, the halo's into pl
    // (p is not needed again: a replacement starts a fresh direction).
    if (a.r0d == nullptr || status > 1) break;
    ++ep;
    if (own) put_f3(a.tdx + 3 * static_cast<size_t>(slot), ep, dx);
    sts4(pl, t, dx);
    auto dx_word = [&](int i, const unsigned long long*& addr, int& dest) {
      const int hh = i / 3, c = i % 3;
      addr = a.tdx + 3 * static_cast<size_t>(a.halo_slot[h0 + hh]) + c;
      dest = static_cast<int>(pl - dyn) + 4 * (kPcgThreads + hh) + c;
    };
    if (!__syncthreads_and(collect_rest(t, 3 * nh, dx_word, ep, dyn, a.abort_flag))) return;
    // rt = r0d - (K + D) dx of the own row, in FP64 (difference form), read
    // from the slot-ELL in memory: this runs once or twice per solve, and the
    // register copy of the row would raise the kernel's register peak.
    f3 rt = {0.0f, 0.0f, 0.0f};
    double e2 = 0.0;
    if (own) {
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int k = 0; k < e_w; ++k) {
        const f3 dn = lds4(pl, a.sell_lcol[e_off + k * kSlice + e_lane]);
        const double v = a.sell_val[e_off + k * kSlice + e_lane];
        ax += v * (static_cast<double>(dn.x) - static_cast<double>(dx.x));
        ay += v * (static_cast<double>(dn.y) - static_cast<double>(dx.y));
        az += v * (static_cast<double>(dn.z) - static_cast<double>(dx.z));
      }
      const double ddd = dd;
      const size_t i3 = 3 * static_cast<size_t>(row);
      const double ex = a.r0d[i3] - (ax + ddd * dx.x);
      const double ey = a.r0d[i3 + 1] - (ay + ddd * dx.y);
      const double ez = a.r0d[i3 + 2] - (az + ddd * dx.z);
      e2 = ex * ex + ey * ey + ez * ez;
      rt = {static_cast<float>(ex), static_cast<float>(ey), static_cast<float>(ez)};
      // The residual of x0 + dxc + dx: kept for a replacement (r0d is not
      // read again once the test passes).
      a.r0d[i3] = ex;
      a.r0d[i3 + 1] = ey;
      a.r0d[i3 + 2] = ez;
    }
    // Exchange 4: the slot wavefronts' sums of e2 (words [24, 32) of each
    // workgroup's partials), totalled by every workgroup in the same order.
    ++ep;
    {
      const double s = wave_sum(e2);
      if (lane0 && !cwg) put_double(my_tp + 3 * kWaveWords, ep, s);
    }
    auto true_word = [&](int i, const unsigned long long*& addr, int& dest) {
      const int w = i / kWaveWords, h = i % kWaveWords;
      addr = a.tp + kPcgPartWords * static_cast<size_t>(S.wg0 + w) + 3 * kWaveWords + h;
      dest = static_cast<int>(prz - dyn) + kWaveWords * w + h;
    };
    if (!__syncthreads_and(
            collect_rest(t, kWaveWords * S.nwg, true_word, ep, dyn, a.abort_flag)))
      return;
    const double true2 = part_total(prz, kPcgWaves * S.nwg);
    true_ratio = ref2 > 0.0 ? true2 / ref2 : 0.0;
    if (true2 <= fmax(abs2, rel2 * ref2)) break;
    if (replaced >= a.max_replace) {
      unverified = 1;
      break;
    }
    ++replaced;
    // Residual replacement: r from the FP64 residual; dx moves into the FP64
    // committed correction and restarts at zero; the search restarts (z = M r
    // and the coarse members' sums follow at the start).
    if (own) {
      r = rt;
      double* c = a.dxc + 3 * static_cast<size_t>(row);
      const double d3v[3] = {dx.x, dx.y, dx.z};
      for (int k = 0; k < 3; ++k) c[k] = (replaced == 1 ? 0.0 : c[k]) + d3v[k];
    }
    dx = {0.0f, 0.0f, 0.0f};
    start = true;
  }
  if (own) {
    if (replaced == 0) {
      st3(a.x, row, ld3(a.x, row) + dx);
      if (a.dx_out != nullptr) st3(a.dx_out, row, dx);
    } else {
      // x0 + dxc + dx rounded once.
      const double* c = a.dxc + 3 * static_cast<size_t>(row);
      const f3 x0 = ld3(a.x, row);
      const double t0 = c[0] + dx.x, t1 = c[1] + dx.y, t2 = c[2] + dx.z;
      st3(a.x, row, f3{static_cast<float>(x0.x + t0), static_cast<float>(x0.y + t1),
                       static_cast<float>(x0.z + t2)});
      if (a.dx_out != nullptr)
        st3(a.dx_out, row,
            f3{static_cast<float>(t0), static_cast<float>(t1), static_cast<float>(t2)});
    }
  }
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
      atomicMax(&a.ctl->pcg_max_solve, iter);
      if (status == 2) atomicAdd(&a.ctl->pcg_stalled, 1);
      if (replaced != 0) atomicAdd(&a.ctl->pcg_restarts, replaced);
      if (unverified != 0) atomicAdd(&a.ctl->pcg_unverified, 1);
      if (true_ratio >= 0.0)
        atomicMax(&a.ctl->max_true_bits,
                  static_cast<unsigned long long>(__double_as_longlong(true_ratio)));
      if (a.restart != 0) {
        atomicAdd(&a.ctl->pcg_restarts, 1);
      } else if (sys == 0) {
        atomicAdd(&a.ctl->pcg_solves, 1);
        a.ctl->last_rel = static_cast<float>(rr0 > 0.0 ? sqrt(rr / rr0) : 0.0);
      }
    }
  }
}

struct Occupancy {
  int blocks_per_cu = -1;
  size_t lds = static_cast<size_t>(-1);
};
Occupancy g_occ[2];
int g_cus = 0;

const void* persistent_fn(int dense) {
  return dense ? reinterpret_cast<const void*>(k_pcg_persistent<2>)
               : reinterpret_cast<const void*>(k_pcg_persistent<1>);
}

}  // namespace

size_t pcg_persistent_lds(const PersistDev& a) {
  const size_t hl = static_cast<size_t>(kPcgThreads) + a.nhalo_max;
  const size_t qs_stride = a.dense ? kQsStride<2> : kQsStride<1>;
  return sizeof(float) * (4 * static_cast<size_t>(a.nc_max + 1) +
                          ((qs_stride * static_cast<size_t>(a.need_max) + 3) & ~size_t{3}) +
                          24 * static_cast<size_t>(a.nwg_max) +
                          4 * static_cast<size_t>(a.ncb_max) + 8 * hl);
}

bool pcg_persistent_fits(int grid, size_t shmem, int dense) {
  if (shmem > static_cast<size_t>(kPcgLdsBytes)) return false;
  Occupancy& oc = g_occ[dense ? 1 : 0];
  if (oc.blocks_per_cu < 0 || oc.lds != shmem) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    if (!prop.cooperativeLaunch) return false;
    const void* fn = persistent_fn(dense);
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(shmem)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    int nb = 0;
    const hipError_t e =
        dense ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pcg_persistent<2>,
                                                             kPcgThreads, shmem)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pcg_persistent<1>,
                                                             kPcgThreads, shmem);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    oc.blocks_per_cu = nb;
    oc.lds = shmem;
    g_cus = prop.multiProcessorCount;
  }
  return grid > 0 && grid <= oc.blocks_per_cu * g_cus;
}

hipError_t launch_pcg_persistent(hipStream_t s, const PersistDev& a,
                                 float rel_tol) {
  PersistDev arg = a;
  float tol = rel_tol;
  void* args[] = {&arg, &tol};
  const size_t shmem = pcg_persistent_lds(a);
  // The grid was sized to fit the device (pcg_persistent_fits: at most the
  // occupancy-limited number of resident workgroups) and every wait is
  // bounded, so a plain launch is used: a cooperative launch left the device
  // idle for ≈11 µs on each side of every solve (≈3 % of the C2 step).
  // SILK_HIP_COOP=1 asks for the cooperative launch and its residency
  // guarantee (a refused one falls back to the plain launch).  Without that
  // guarantee a failure is reported, never a hang.
  static const bool coop_env = [] {
    const char* e = std::getenv("SILK_HIP_COOP");
    return e != nullptr && e[0] == '1';
  }();
  if (coop_env) {
    const hipError_t e = hipLaunchCooperativeKernel(persistent_fn(a.dense), dim3(a.grid),
                                                    dim3(kPcgThreads), args,
                                                    static_cast<unsigned>(shmem), s);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
  }
  if (a.dense)
    k_pcg_persistent<2><<<dim3(a.grid), dim3(kPcgThreads), static_cast<unsigned>(shmem), s>>>(
        arg, tol);
  else
    k_pcg_persistent<1><<<dim3(a.grid), dim3(kPcgThreads), static_cast<unsigned>(shmem), s>>>(
        arg, tol);
  return hipGetLastError();
}

}  // namespace silk::amd
