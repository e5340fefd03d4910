// Contact colouring on the device (replaces greedy_color / resolve_conflicts
// and the colour sort of contact_coloring.cu:69-373).
//
// Contacts conflict when they share a movable vertex (a slot with a contact
// vertex and a non-zero gamma).  Each call colours the contacts appended
// since the last one, [n_old, n); earlier contacts keep their colours.  The
// colouring is first-fit in contact order: contact c takes the smallest
// colour >= 1 that no conflicting contact of lower index holds.  It is
// computed in Jones-Plassmann rounds with the contact index as priority: a
// contact is coloured in the round in which it is the lowest uncoloured
// contact at every one of its vertices, so every conflicting contact of
// lower index is coloured before it and the result equals the sequential
// first-fit colouring.  Contacts coloured in one round share no vertex, so
// their colour masks are updated without conflicts.  The reference's
// speculative colouring depends on thread timing; this one is deterministic.
//
// The same workgroup then sorts all n contacts by colour (stably, so ids
// ascend inside a colour) into the layout the contact update reads: one
// contiguous range per colour and the ranges' offsets.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>

#include "hip/collision_kernels.hpp"

namespace silk::amd {
namespace {

constexpr int kColorThreads = 1024;
constexpr int kColorWaves = kColorThreads / 64;
constexpr int kColoredNow = 1 << 30;  // marks a colour written in this round

__device__ __forceinline__ bool movable(const ContactDev& c, int s) {
  return c.state[s] >= 0 && c.gamma[s] != 0.0f;
}

// Static LDS of the kernel; the rest of the workgroup's LDS holds the
// vertex masks and minima when they fit.
constexpr size_t kColorStaticLds =
    4 * (4 + (kMaxContactColors + 2) + kColorWaves * (kMaxContactColors + 1));
constexpr size_t kColorLdsBytes = 160 * 1024 - kColorStaticLds - 1024;

// Stores made before it are visible to every thread of the workgroup after
// it.  With the masks in global memory the fence makes the stores that went
// through the L1 cache, and the atomics at L2, visible to the loads after
// the barrier.
template <bool kLds>
__device__ __forceinline__ void phase_end() {
  if (!kLds) __threadfence();
  __syncthreads();
}

// kLds: the per-vertex colour masks and minima in LDS (nact * (kColorWords
// + 1) words), else in the global scratch of ColorArgs.
template <bool kLds>
__global__ void __launch_bounds__(kColorThreads) k_color_contacts(ColorArgs a) {
  __shared__ int s_left, s_over, s_maxc, s_rounds;
  __shared__ int s_start[kMaxContactColors + 2];
  __shared__ int s_wave[kColorWaves][kMaxContactColors + 1];
  extern __shared__ unsigned s_dyn[];
  if (kLds) {
    a.mask = s_dyn;
    a.head = reinterpret_cast<int*>(s_dyn + static_cast<size_t>(a.nact) * kColorWords);
  }
  const int t = static_cast<int>(threadIdx.x);
  const int nw = a.nact * kColorWords;
  for (int i = t; i < nw; i += kColorThreads) a.mask[i] = 0u;
  for (int i = t; i < a.nact; i += kColorThreads) a.head[i] = INT_MAX;
  for (int c = a.n_old + t; c < a.n; c += kColorThreads) a.color[c] = 0;
  for (int k = t; k < kMaxContactColors + 2; k += kColorThreads) s_start[k] = 0;
  if (t == 0) {
    s_left = a.n - a.n_old;
    s_over = 0;
    s_maxc = 0;
    s_rounds = 0;
  }
  phase_end<kLds>();
  // Colours of the earlier contacts into their vertices' masks.
  for (int c = t; c < a.n_old; c += kColorThreads) {
    const ContactDev& cc = a.cin[c];
    const int k = a.color[c];
    atomicMax(&s_maxc, k);
    for (int s = 0; s < 4; ++s)
      if (movable(cc, s))
        atomicOr(&a.mask[cc.state[s] * kColorWords + (k >> 5)], 1u << (k & 31));
  }
  phase_end<kLds>();
  // Every round colours at least the lowest uncoloured contact, so n - n_old
  // rounds bound the loop.
  for (int round = 0; round <= a.n - a.n_old; ++round) {
    if (s_left == 0) break;  // read after a barrier: the same in every wave
    for (int c = a.n_old + t; c < a.n; c += kColorThreads) {
      if (a.color[c] != 0) continue;
      const ContactDev& cc = a.cin[c];
      for (int s = 0; s < 4; ++s)
        if (movable(cc, s)) atomicMin(&a.head[cc.state[s]], c);
    }
    phase_end<kLds>();
    for (int c = a.n_old + t; c < a.n; c += kColorThreads) {
      if (a.color[c] != 0) continue;
      const ContactDev& cc = a.cin[c];
      bool ready = true;
      for (int s = 0; s < 4; ++s)
        if (movable(cc, s) && a.head[cc.state[s]] != c) ready = false;
      if (!ready) continue;
      int k = 0;
      for (int w = 0; w < kColorWords && k == 0; ++w) {
        unsigned used = w == 0 ? 1u : 0u;  // colour 0 is not a colour
        for (int s = 0; s < 4; ++s)
          if (movable(cc, s)) used |= a.mask[cc.state[s] * kColorWords + w];
        if (used != 0xffffffffu) k = 32 * w + __builtin_ctz(~used);
      }
      if (k == 0 || k > kMaxContactColors) {
        s_over = 1;
        k = kMaxContactColors;
      }
      for (int s = 0; s < 4; ++s)
        if (movable(cc, s))
          atomicOr(&a.mask[cc.state[s] * kColorWords + (k >> 5)], 1u << (k & 31));
      a.color[c] = k | kColoredNow;
      atomicMax(&s_maxc, k);
      atomicSub(&s_left, 1);
    }
    phase_end<kLds>();
    // The contacts that were uncoloured at the round's start release their
    // vertices for the next round's minima.
    for (int c = a.n_old + t; c < a.n; c += kColorThreads) {
      const int col = a.color[c];
      if (col != 0 && !(col & kColoredNow)) continue;
      const ContactDev& cc = a.cin[c];
      for (int s = 0; s < 4; ++s)
        if (movable(cc, s)) a.head[cc.state[s]] = INT_MAX;
      if (col & kColoredNow) a.color[c] = col & ~kColoredNow;
    }
    if (t == 0) s_rounds = round + 1;
    phase_end<kLds>();
  }
  if (s_left != 0) {  // not reached: every round colours a contact
    if (t == 0) {
      a.meta[0] = 0;
      a.meta[1] = s_rounds;
      a.meta[2] = 1;
    }
    return;
  }
  // Stable counting sort by colour: ranges in colour order, ids ascending.
  const int ncol = s_maxc;
  for (int c = t; c < a.n; c += kColorThreads) atomicAdd(&s_start[a.color[c]], 1);
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int k = 1; k <= ncol; ++k) {
      const int cnt = s_start[k];
      s_start[k] = run;
      run += cnt;
    }
  }
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  for (int base = 0; base < a.n; base += kColorThreads) {
    for (int i = t; i < kColorWaves * (kMaxContactColors + 1); i += kColorThreads)
      (&s_wave[0][0])[i] = 0;
    __syncthreads();
    const int c = base + t;
    const bool valid = c < a.n;
    const int k = valid ? a.color[c] : 0;
    // Rank among the wave's lanes of the same colour, and the wave's count
    // per colour.
    int rank = 0;
    unsigned long long pending = __ballot(valid);
    while (pending != 0ull) {
      const int leader = __ffsll(static_cast<long long>(pending)) - 1;
      const int kk = __shfl(k, leader);
      const unsigned long long m = __ballot(valid && k == kk);
      if (valid && k == kk) rank = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == leader) s_wave[wave][kk] = __popcll(m);
      pending &= ~m;
    }
    __syncthreads();
    // Per colour: the waves' bases in wave order, continuing the colour's
    // running start.
    for (int kk = 1 + t; kk <= ncol; kk += kColorThreads) {
      int run = s_start[kk];
      for (int w = 0; w < kColorWaves; ++w) {
        const int cnt = s_wave[w][kk];
        s_wave[w][kk] = run;
        run += cnt;
      }
      s_start[kk] = run;
    }
    __syncthreads();
    if (valid) {
      ContactDev out = a.cin[c];
      out.id = c;
      a.cout[s_wave[wave][k] + rank] = out;
    }
    __syncthreads();
  }
  // s_start[k] is now the end of colour k's range.
  if (t == 0) {
    a.off[0] = 0;
    for (int k = 1; k <= ncol; ++k) a.off[k] = s_start[k];
    a.meta[0] = ncol;
    a.meta[1] = s_rounds;
    a.meta[2] = s_over;
  }
}

// Counting sort of the colour-ordered contacts by (group, colour).  Inside
// one (group, colour) range the order is arbitrary: contacts of one colour
// share no movable vertex, so the update's result does not depend on it.
__global__ void __launch_bounds__(kColorThreads) k_group_contacts(GroupArgs a) {
  __shared__ int s_part[kColorThreads];
  const int t = static_cast<int>(threadIdx.x);
  const int ncol = a.meta[0];
  if (ncol <= 0) return;  // the colouring failed; the caller reports it
  const int nkey = a.ngroup * ncol;
  for (int k = t; k < nkey; k += kColorThreads) a.count[k] = 0;
  phase_end<false>();
  for (int i = t; i < a.n; i += kColorThreads) {
    const int id = a.cin[i].id;
    atomicAdd(&a.count[a.group[id] * ncol + a.color[id] - 1], 1);
  }
  phase_end<false>();
  // Exclusive scan of the counts: a chunk per thread, the chunk totals
  // scanned in LDS.
  const int chunk = (nkey + kColorThreads - 1) / kColorThreads;
  const int k0 = min(nkey, t * chunk), k1 = min(nkey, k0 + chunk);
  int sum = 0;
  for (int k = k0; k < k1; ++k) sum += a.count[k];
  s_part[t] = sum;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int j = 0; j < kColorThreads; ++j) {
      const int v = s_part[j];
      s_part[j] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = s_part[t];
  for (int k = k0; k < k1; ++k) {
    const int v = a.count[k];
    a.off[k] = run;
    a.count[k] = run;  // cursor
    run += v;
  }
  if (t == 0) a.off[nkey] = a.n;
  phase_end<false>();
  for (int i = t; i < a.n; i += kColorThreads) {
    const ContactDev c = a.cin[i];
    const int pos = atomicAdd(&a.count[a.group[c.id] * ncol + a.color[c.id] - 1], 1);
    a.cout[pos] = c;
  }
}

}  // namespace

hipError_t launch_group_contacts(hipStream_t s, const GroupArgs& a) {
  if (a.n <= 0 || a.ngroup <= 0) return hipErrorInvalidValue;
  k_group_contacts<<<1, kColorThreads, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_color_contacts(hipStream_t s, const ColorArgs& a) {
  if (a.n <= 0 || a.n_old < 0 || a.n_old > a.n || a.nact <= 0) return hipErrorInvalidValue;
  const size_t lds = static_cast<size_t>(a.nact) * (kColorWords + 1) * sizeof(unsigned);
  // SILK_HIP_COLOR_GLOBAL=1 keeps the masks in global memory (tests).
  const char* env = std::getenv("SILK_HIP_COLOR_GLOBAL");
  const bool force_global = env != nullptr && env[0] == '1';
  if (lds <= kColorLdsBytes && !force_global) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void*>(k_color_contacts<true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kColorLdsBytes));
    if (attr == hipSuccess) {
      k_color_contacts<true><<<1, kColorThreads, lds, s>>>(a);
      return hipGetLastError();
    }
    (void)hipGetLastError();
  }
  k_color_contacts<false><<<1, kColorThreads, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace silk::amd
