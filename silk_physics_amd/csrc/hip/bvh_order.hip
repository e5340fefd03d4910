// Per-step Morton re-ordering of the broadphase trees' primitives and of the
// query visiting orders on the device.  The reference recomputes the Morton
// codes of its OIBVH primitives and radix-sorts them on every update
// (oibvh.cuh:521-560,603-633); the trees here keep their implicit layout and
// only their primitive order changes, so a refit afterwards rebuilds them.
//
// Codes are those of the host order made at set-up (10 bits per axis of the
// item centroid over the segment's centroid bounds, ties in index order: the
// radix sort is stable), so the first re-ordering at the set-up positions
// reproduces the host order exactly.  All segments are sorted together on
// (segment << 32 | code): segments are contiguous in key order, so every
// segment keeps its place and is sorted within it.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "hip/collision_kernels.hpp"
#include "hip/device_common.hpp"

namespace silk::amd {
namespace {

constexpr int kOrderBlock = 256;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Order-preserving map of a float to an unsigned integer (and back), so
// that bounds can be reduced with integer atomics.
__device__ __forceinline__ uint32_t enc(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float dec(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ void centroid(const OrderSeg& s, const float* p, int local,
                                         float c[3]) {
  for (int d = 0; d < 3; ++d) {
    float acc = 0.0f;
    for (int k = 0; k < s.nvert; ++k) acc += p[3 * s.src[s.nvert * local + k] + d];
    c[d] = acc / static_cast<float>(s.nvert);
  }
}

// Empty bounds: encodings of +inf (minima) and -inf (maxima).
__global__ void k_seg_bounds_init(uint32_t* box, int nseg) {
  const int i = blockIdx.x * kOrderBlock + threadIdx.x;
  if (i < 6 * nseg) box[i] = (i % 6) < 3 ? 0xff800000u : 0x007fffffu;
}

// Workgroup b reduces the centroids of up to kOrderItemsPerBlock items of one
// segment (blk[b] = (segment, first item)) and folds them into the segment's
// bounds with one atomic per coordinate and side.
__global__ void __launch_bounds__(kOrderBlock) k_seg_bounds(const OrderSeg* segs,
                                                            const int2* blk,
                                                            const float* p, uint32_t* box) {
  __shared__ float red[6][kOrderBlock];
  const int2 bs = blk[blockIdx.x];
  const OrderSeg s = segs[bs.x];
  const int end = min(s.n, bs.y + kOrderItemsPerBlock);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = bs.y + static_cast<int>(threadIdx.x); i < end; i += kOrderBlock) {
    float c[3];
    centroid(s, p, i, c);
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], c[d]);
      hi[d] = fmaxf(hi[d], c[d]);
    }
  }
  for (int d = 0; d < 3; ++d) {
    red[d][threadIdx.x] = lo[d];
    red[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int w = kOrderBlock / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + w]);
        red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3)
    atomicMin(&box[6 * bs.x + threadIdx.x], enc(red[threadIdx.x][0]));
  else if (threadIdx.x < 6)
    atomicMax(&box[6 * bs.x + threadIdx.x], enc(red[threadIdx.x][0]));
}

__global__ void __launch_bounds__(kOrderBlock) k_morton(const OrderSeg* segs,
                                                        const int* seg_of, int total,
                                                        const float* p, const uint32_t* box,
                                                        unsigned long long* keys, int* vals) {
  const int i = blockIdx.x * kOrderBlock + threadIdx.x;
  if (i >= total) return;
  const int si = seg_of[i];
  const OrderSeg s = segs[si];
  const int local = i - s.off;
  float c[3];
  centroid(s, p, local, c);
  uint32_t code = 0;
  for (int d = 0; d < 3; ++d) {
    const float lo = dec(box[6 * si + d]), ext = dec(box[6 * si + 3 + d]) - lo;
    const float t = ext > 0.0f ? (c[d] - lo) / ext : 0.0f;
    const uint32_t q = static_cast<uint32_t>(fminf(fmaxf(t * 1023.0f, 0.0f), 1023.0f));
    code |= spread10(q) << (2 - d);
  }
  keys[i] = (static_cast<unsigned long long>(si) << 32) | code;
  vals[i] = local;
}

__global__ void __launch_bounds__(kOrderBlock) k_apply_order(const OrderSeg* segs,
                                                             const int* seg_of, int total,
                                                             const int* vals) {
  const int i = blockIdx.x * kOrderBlock + threadIdx.x;
  if (i >= total) return;
  const OrderSeg s = segs[seg_of[i]];
  const int k = i - s.off;
  const int local = vals[i];
  if (s.out_order != nullptr) {
    s.out_order[k] = local;
    return;
  }
  for (int v = 0; v < s.nvert; ++v) s.out_prim[s.nvert * k + v] = s.src[s.nvert * local + v];
  s.out_id[k] = s.gid0 + local;
}

int key_end_bit(int nseg) {
  int bits = 0;
  while ((1 << bits) < nseg) ++bits;
  return 32 + bits;
}

}  // namespace

size_t order_sort_temp_bytes(int total, int nseg) {
  size_t bytes = 0;
  if (total <= 0 || nseg <= 0) return 0;
  if (hipcub::DeviceRadixSort::SortPairs(
          nullptr, bytes, static_cast<const unsigned long long*>(nullptr),
          static_cast<unsigned long long*>(nullptr), static_cast<const int*>(nullptr),
          static_cast<int*>(nullptr), total, 0, key_end_bit(nseg), nullptr) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t launch_order_rebuild(hipStream_t s, const float* p, const OrderSeg* segs, int nseg,
                                const int* seg_off, const int* seg_of, int total,
                                OrderScratch sc) {
  (void)seg_off;
  if (total <= 0 || nseg <= 0 || sc.nblk <= 0) return hipSuccess;
  k_seg_bounds_init<<<(6 * nseg + kOrderBlock - 1) / kOrderBlock, kOrderBlock, 0, s>>>(sc.box,
                                                                                      nseg);
  k_seg_bounds<<<sc.nblk, kOrderBlock, 0, s>>>(segs, sc.blk, p, sc.box);
  const int grid = (total + kOrderBlock - 1) / kOrderBlock;
  k_morton<<<grid, kOrderBlock, 0, s>>>(segs, seg_of, total, p, sc.box, sc.keys_in,
                                        sc.vals_in);
  size_t bytes = sc.temp_bytes;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      sc.temp, bytes, sc.keys_in, sc.keys_out, sc.vals_in, sc.vals_out, total, 0,
      key_end_bit(nseg), s);
  if (e != hipSuccess) return e;
  k_apply_order<<<grid, kOrderBlock, 0, s>>>(segs, seg_of, total, sc.vals_out);
  return hipGetLastError();
}

}  // namespace silk::amd
