// Collision kernels of the step.  See collision_kernels.hpp.
//
// Reference: collision/object_collider.cu:378-475 (swept, padded boxes),
// oibvh.cuh:417-516 (stack traversal, closed-interval overlap, self pairs
// once), find_collision.cu:45-92 (filters), dcd.cu:23-315 (DCD),
// cpu/collision/interval_root_finder.cpp + ccd.cpp (the CCD restated here in
// FP64), contact_constraints.cu:203-411 (Daviet Gauss-Seidel).
#include <algorithm>
#include <cfloat>

#include "hip/collision_kernels.hpp"
#include "hip/device_common.hpp"

namespace silk::amd {

namespace {

constexpr int kRefitThreads = 1024;

__device__ __forceinline__ void swept_box(const CollVerts& v, const int* gv,
                                          int n, float pad, Box6& b) {
  for (int d = 0; d < 3; ++d) {
    float mn = 1e30f, mx = -1e30f;
    for (int k = 0; k < n; ++k) {
      const float a = v.p0[3 * gv[k] + d];
      const float c = v.p1[3 * gv[k] + d];
      mn = fminf(mn, fminf(a, c));
      mx = fmaxf(mx, fmaxf(a, c));
    }
    b.mn[d] = mn - pad;
    b.mx[d] = mx + pad;
  }
}

__device__ __forceinline__ bool overlap(const Box6& a, const Box6& b) {
  for (int d = 0; d < 3; ++d) {
    const float lo = fmaxf(a.mn[d], b.mn[d]);
    const float hi = fminf(a.mx[d], b.mx[d]);
    if (!(lo <= hi)) return false;
  }
  return true;
}

__device__ __forceinline__ Box6 merge(const Box6& a, const Box6& b) {
  Box6 r;
  for (int d = 0; d < 3; ++d) {
    r.mn[d] = fminf(a.mn[d], b.mn[d]);
    r.mx[d] = fmaxf(a.mx[d], b.mx[d]);
  }
  return r;
}

__device__ __forceinline__ Box6 empty_box() {
  Box6 r;
  for (int d = 0; d < 3; ++d) {
    r.mn[d] = FLT_MAX;
    r.mx[d] = -FLT_MAX;
  }
  return r;
}

// ---------------------------------------------------------------- refit
// Refit in two launches.  Leaves: workgroup b of tree t (blocks
// off[t] <= b < off[t + 1]) computes the primitive boxes and leaf boxes of
// 256 consecutive leaves and the levels of their 256-leaf subtree in LDS
// (boxes merge with min/max, so the order of merging does not matter).
constexpr int kRefitLeaves = 256;

__global__ void __launch_bounds__(kRefitLeaves)
    k_refit_leaves(CollVerts v, const TreeDev* __restrict__ trees,
                   const int* __restrict__ off, int ntrees) {
  __shared__ Box6 lb[kRefitLeaves];
  int ti = 0;
  while (ti + 1 < ntrees && off[ti + 1] <= static_cast<int>(blockIdx.x)) ++ti;
  const TreeDev t = trees[ti];
  const int tid = static_cast<int>(threadIdx.x);
  const int l0 = (static_cast<int>(blockIdx.x) - off[ti]) * kRefitLeaves;
  const int l = l0 + tid;
  Box6 b = empty_box();
  if (l < t.P) {
    const int j1 = min(l * kLeafPrims + kLeafPrims, t.nprim);
    for (int j = l * kLeafPrims; j < j1; ++j) {
      int gv[3];
      for (int k = 0; k < t.nvert; ++k) gv[k] = t.prim[t.nvert * j + k];
      Box6 pb;
      swept_box(v, gv, t.nvert, t.pad, pb);
      t.sbox[j] = pb;
      b = merge(b, pb);
    }
    t.node[t.P + l] = b;
  }
  lb[tid] = b;
  __syncthreads();
  for (int k = 1; (kRefitLeaves >> k) >= 1 && (t.P >> k) >= 1; ++k) {
    const int w = min(kRefitLeaves >> k, t.P >> k);
    Box6 m;
    if (tid < w) m = merge(lb[2 * tid], lb[2 * tid + 1]);
    __syncthreads();
    if (tid < w) {
      lb[tid] = m;
      t.node[((t.P + l0) >> k) + tid] = m;
    }
    __syncthreads();
  }
}

// Refit, top: the levels above the 256-leaf subtrees, one workgroup per tree.
__global__ void __launch_bounds__(kRefitThreads)
    k_refit_top(const TreeDev* __restrict__ trees) {
  const TreeDev t = trees[blockIdx.x];
  for (int half = (t.P / kRefitLeaves) >> 1; half >= 1; half >>= 1) {
    for (int j = half + threadIdx.x; j < 2 * half; j += blockDim.x)
      t.node[j] = merge(t.node[2 * j], t.node[2 * j + 1]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- traversal
constexpr int kTraverseThreads = kTraverseWg;
constexpr int kPairStage = 2048;  // pairs a workgroup stages in LDS
// A subtree a workgroup stages in LDS: at most this many nodes and leaf
// primitive boxes (about 40 KB of LDS with the pair stage, four workgroups
// per CU; twice these sizes halved the residency and lost more than the
// larger trees gained).
constexpr int kSubNodes = 511;
constexpr int kSubPrims = 512;

__device__ __forceinline__ Box6 lds_box(const float* b, int i) {
  Box6 r;
  for (int d = 0; d < 3; ++d) {
    r.mn[d] = b[6 * i + d];
    r.mx[d] = b[6 * i + 3 + d];
  }
  return r;
}

// All passes of a broadphase in one launch.  A work item is (job, subtree,
// query): lane `item` walks the subtree rooted at heap node 2^depth + sub
// for the query at position li of the job's Morton order.  The subtree walk
// is stackless on the implicit heap (descend to the left child on overlap,
// otherwise climb while on a right child and step to the right sibling).
// Parent boxes contain their children, so the leaves reached over all
// subtrees are the leaves of a walk from the root (oibvh.cuh:417-516).
// Pairs are staged in LDS and written out with one counter reservation per
// workgroup (one shared counter per pair serialised the whole launch on it);
// pairs beyond the stage go out one by one.  When all items of a workgroup
// walk the same subtree (consecutive items are consecutive queries of one
// subtree) and it fits, its node boxes and leaf primitive boxes are first
// copied to LDS with coalesced loads, and the walks read them there.
__global__ void __launch_bounds__(kTraverseThreads)
    k_broadphase(CollVerts v, const TraverseJob* __restrict__ jobs, int njobs,
                 PairOut pt, PairOut ee) {
  __shared__ int2 stage[kPairStage];
  __shared__ int stage_n, stage_base;
  __shared__ float snode[6 * kSubNodes];
  __shared__ float sprim[6 * kSubPrims];
  // Job ranges are whole workgroups: the job is uniform over the workgroup.
  const int wg0 = blockIdx.x * blockDim.x;
  int j = 0;
  while (j + 1 < njobs && jobs[j + 1].item0 <= wg0) ++j;
  j = __builtin_amdgcn_readfirstlane(j);
  const TraverseJob& job = jobs[j];
  const int local = wg0 + static_cast<int>(threadIdx.x) - job.item0;
  if (threadIdx.x == 0) stage_n = 0;
  __syncthreads();
  const PairOut out = job.kind == 0 ? pt : ee;
  bool active = local < job.nitem;
  if (job.gate_a != nullptr && job.gate_b != nullptr &&
      !overlap(*job.gate_a, *job.gate_b))
    active = false;
  // The subtree of this workgroup's items, when it is one and fits: heap
  // node r at depth `depth`, levels 0..D below it, leaves' primitives
  // [jb, je).
  const TreeDev& tj = job.tree;
  const int first = wg0 - job.item0;
  const int last = min(first + kTraverseThreads, job.nitem) - 1;
  const int levels = 31 - __clz(tj.P) - job.depth;  // D
  const int sub_nodes = (2 << levels) - 1;
  const int r0 = (1 << job.depth) + (first >= 0 && job.q.n > 0 ? first / job.q.n : 0);
  const int jb = ((r0 << levels) - tj.P) * kLeafPrims;
  const int je = min(jb + (kLeafPrims << levels), tj.nprim);
  const bool cached = first >= 0 && last >= first && job.q.n > 0 &&
                      first / job.q.n == last / job.q.n && levels >= 0 &&
                      sub_nodes <= kSubNodes && (kLeafPrims << levels) <= kSubPrims;
  if (cached) {
    for (int i = threadIdx.x; i < 6 * sub_nodes; i += kTraverseThreads) {
      const int li = i / 6, c = i - 6 * li;
      const int k = 31 - __clz(li + 1);
      const int heap = (r0 << k) + (li + 1 - (1 << k));
      snode[i] = reinterpret_cast<const float*>(tj.node)[6 * heap + c];
    }
    for (int i = threadIdx.x; i < 6 * max(0, je - jb); i += kTraverseThreads)
      sprim[i] = reinterpret_cast<const float*>(tj.sbox)[6 * jb + i];
    __syncthreads();
  }
  if (active) {
    const QuerySet q = job.q;
    const TreeDev t = job.tree;
    const int kind = job.kind, self = job.self;
    const int sub = local / q.n;
    const int li = q.order[local - sub * q.n];
    const int qi = q.gid0 + li;
    int qv[2];
    for (int k = 0; k < q.nvert; ++k) qv[k] = q.prim[q.nvert * li + k];
    Box6 qb;
    swept_box(v, qv, q.nvert, q.pad, qb);
    float qinv = 0.0f;
    for (int k = 0; k < q.nvert; ++k) qinv += v.invm[qv[k]];
    const int r = (1 << job.depth) + sub;
    int node = r;
    for (;;) {
      bool down = false;
      Box6 nb;
      if (cached) {
        const int k = (31 - __clz(node)) - job.depth;
        nb = lds_box(snode, (1 << k) - 1 + (node - (r << k)));
      } else {
        nb = t.node[node];
      }
      if (overlap(qb, nb)) {
        if (node < t.P) {
          down = true;
        } else {
          const int j0 = (node - t.P) * kLeafPrims;
          const int j1 = min(j0 + kLeafPrims, t.nprim);
          for (int jj = j0; jj < j1; ++jj) {
            if (!overlap(qb, cached ? lds_box(sprim, jj - jb) : t.sbox[jj])) continue;
            int tv[3];
            float inv = qinv;
            for (int k = 0; k < t.nvert; ++k) {
              tv[k] = t.prim[t.nvert * jj + k];
              inv += v.invm[tv[k]];
            }
            if (inv == 0.0f) continue;  // all pinned / kinematic
            const int tid_orig = t.prim_id[jj];
            if (self) {
              bool adjacent = false;
              for (int a = 0; a < t.nvert; ++a)
                for (int b = 0; b < q.nvert; ++b) adjacent |= (tv[a] == qv[b]);
              if (adjacent) continue;
              if (kind == 1 && !(tid_orig < qi)) continue;  // each pair once
            }
            const int2 pr = make_int2(tid_orig, qi);
            const int slot = atomicAdd(&stage_n, 1);
            if (slot < kPairStage) {
              stage[slot] = pr;
            } else {
              const int g = atomicAdd(out.count, 1);
              if (g < out.cap) out.pairs[g] = pr;
            }
          }
        }
      }
      if (down) {
        node <<= 1;
        continue;
      }
      while (node != r && (node & 1)) node >>= 1;
      if (node == r) break;
      ++node;
    }
  }
  __syncthreads();
  const int staged = min(stage_n, kPairStage);
  if (threadIdx.x == 0) stage_base = staged > 0 ? atomicAdd(out.count, staged) : 0;
  __syncthreads();
  for (int i = threadIdx.x; i < staged; i += blockDim.x)
    if (stage_base + i < out.cap) out.pairs[stage_base + i] = stage[i];
}

// ---------------------------------------------------------------- DCD
__device__ __forceinline__ dev::f3 pos_at(const CollVerts& v, int gv,
                                          float time) {
  const dev::f3 a = dev::ld3(v.p0, gv), b = dev::ld3(v.p1, gv);
  return (1.0f - time) * a + time * b;
}

// exact_pt_uv (dcd.cu:321-351)
__device__ bool pt_uv(dev::f3 x0, dev::f3 x1, dev::f3 x2, dev::f3 x3, float eps,
                      float* u, float* v) {
  using namespace dev;
  const f3 x21 = x2 - x1, x31 = x3 - x1, x01 = x0 - x1;
  const float a = dot3(x21, x21), b = dot3(x21, x31), c = dot3(x31, x31);
  const float d = dot3(x21, x01), e = dot3(x31, x01);
  const float det = a * c - b * b;
  const float area_eps = eps * fmaxf(a, c);
  if (det < area_eps * area_eps) return false;
  const float b1 = (c * d - b * e) / det;
  const float b2 = (a * e - b * d) / det;
  if (b1 < 0.0f || b2 < 0.0f || b1 + b2 > 1.0f) return false;
  *u = b1;
  *v = b2;
  return true;
}

// exact_ee_uv (dcd.cu:353-394), double inside.
__device__ bool ee_uv(dev::f3 x0, dev::f3 x1, dev::f3 x2, dev::f3 x3, float eps,
                      float* uo, float* vo) {
  const double d1[3] = {double(x1.x) - x0.x, double(x1.y) - x0.y, double(x1.z) - x0.z};
  const double d2[3] = {double(x3.x) - x2.x, double(x3.y) - x2.y, double(x3.z) - x2.z};
  const double r[3] = {double(x0.x) - x2.x, double(x0.y) - x2.y, double(x0.z) - x2.z};
  const double a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  const double b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
  const double c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
  const double e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  const double f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
  const double eps2 = static_cast<double>(eps) * eps;
  const double ls2 = fmax(a, e);
  if (a <= eps2 * ls2 || e <= eps2 * ls2) return false;
  const double den = a * e - b * b;
  double u = den > eps2 * a * e ? fmin(fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
  double v = (b * u + f) / e;
  if (v < 0.0) {
    v = 0.0;
    u = fmin(fmax(-c / a, 0.0), 1.0);
  } else if (v > 1.0) {
    v = 1.0;
    u = fmin(fmax((b - c) / a, 0.0), 1.0);
  }
  *uo = static_cast<float>(u);
  *vo = static_cast<float>(v);
  return true;
}

__device__ void emit(DcdRec* out, int* count, int cap, const DcdRec& r) {
  const int slot = atomicAdd(count, 1);
  if (slot < cap) out[slot] = r;
}

__device__ __forceinline__ float pair_friction(const PrimTables& pt, int oa,
                                               int ob) {
  return static_cast<float>(sqrt(static_cast<double>(pt.obj_friction[oa]) *
                                 static_cast<double>(pt.obj_friction[ob])));
}

__device__ void finish_rec(const CollVerts& v, DcdRec& r, const dev::f3 x[4],
                           dev::f3 disp, float dist) {
  using namespace dev;
  const f3 n = (-1.0f / dist) * disp;
  r.normal[0] = n.x;
  r.normal[1] = n.y;
  r.normal[2] = n.z;
  for (int k = 0; k < 4; ++k) {
    r.x[k][0] = x[k].x;
    r.x[k][1] = x[k].y;
    r.x[k][2] = x[k].z;
  }
  // Prescribed obstacle displacement coef_k (x_k(t1) - x_k(time))
  // (dcd.cu:164-174,245-255).
  f3 pres = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < 4; ++k) {
    if (r.gv[k] < v.nvc) continue;
    pres = pres + r.coef[k] * (ld3(v.p1, r.gv[k]) - x[k]);
  }
  r.presc[0] = pres.x;
  r.presc[1] = pres.y;
  r.presc[2] = pres.z;
}

__global__ void __launch_bounds__(256)
    k_dcd(CollVerts v, PrimTables pt, const int2* __restrict__ ptp,
          const int* __restrict__ ptc, int pt_cap, const int2* __restrict__ eep,
          const int* __restrict__ eec, int ee_cap, float time, float dhat,
          DcdRec* out, int* out_count, int out_cap) {
  using namespace dev;
  const int npt = min(*ptc, pt_cap);
  const int nee = min(*eec, ee_cap);
  const float eps = 1e-6f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npt + nee;
       i += gridDim.x * blockDim.x) {
    DcdRec r;
    f3 x[4];
    float param[4];
    f3 disp;
    if (i < npt) {
      const int2 pr = ptp[i];
      const int t = pr.x, p = pr.y;
      r.type = 0;
      r.gv[0] = pt.point[p];
      r.gv[1] = pt.tri[3 * t];
      r.gv[2] = pt.tri[3 * t + 1];
      r.gv[3] = pt.tri[3 * t + 2];
      for (int k = 0; k < 4; ++k) x[k] = pos_at(v, r.gv[k], time);
      float u, w;
      if (!pt_uv(x[0], x[1], x[2], x[3], eps, &u, &w)) continue;
      const f3 tp = ((1.0f - u - w) * x[1] + u * x[2]) + w * x[3];
      disp = tp - x[0];
      param[0] = 1.0f;
      param[1] = 1.0f - u - w;
      param[2] = u;
      param[3] = w;
      r.coef[0] = 1.0f;
      r.coef[1] = -param[1];
      r.coef[2] = -param[2];
      r.coef[3] = -param[3];
      r.friction = pair_friction(pt, pt.tri_obj[t], pt.point_obj[p]);
    } else {
      const int2 pr = eep[i - npt];
      int a0 = pt.edge[2 * pr.x], a1 = pt.edge[2 * pr.x + 1];
      int b0 = pt.edge[2 * pr.y], b1 = pt.edge[2 * pr.y + 1];
      int oa = pt.edge_obj[pr.x], ob = pt.edge_obj[pr.y];
      // Canonical order: smaller (min, max) vertex pair is edge a.
      const int ka0 = min(a0, a1), ka1 = max(a0, a1);
      const int kb0 = min(b0, b1), kb1 = max(b0, b1);
      if (kb0 < ka0 || (kb0 == ka0 && kb1 < ka1)) {
        int t0 = a0, t1 = a1, to = oa;
        a0 = b0; a1 = b1; oa = ob;
        b0 = t0; b1 = t1; ob = to;
      }
      r.type = 1;
      r.gv[0] = a0;
      r.gv[1] = a1;
      r.gv[2] = b0;
      r.gv[3] = b1;
      for (int k = 0; k < 4; ++k) x[k] = pos_at(v, r.gv[k], time);
      float u, w;
      if (!ee_uv(x[0], x[1], x[2], x[3], eps, &u, &w)) continue;
      const f3 pa = (1.0f - u) * x[0] + u * x[1];
      const f3 pb = (1.0f - w) * x[2] + w * x[3];
      disp = pb - pa;
      param[0] = 1.0f - u;
      param[1] = u;
      param[2] = 1.0f - w;
      param[3] = w;
      r.coef[0] = param[0];
      r.coef[1] = param[1];
      r.coef[2] = -param[2];
      r.coef[3] = -param[3];
      r.friction = pair_friction(pt, oa, ob);
    }
    const float d2 = dot3(disp, disp);
    if (d2 == 0.0f) continue;
    const float dist = sqrtf(d2);
    if (dist >= dhat) continue;
    float den = 0.0f;
    for (int k = 0; k < 4; ++k) {
      r.invm[k] = v.invm[r.gv[k]];
      den += param[k] * param[k] * r.invm[k];
    }
    if (den == 0.0f) continue;
    finish_rec(v, r, x, disp, dist);
    emit(out, out_count, out_cap, r);
  }
}

// ---------------------------------------------------------------- CCD (FP64)
//
// Bit-for-bit restatement of the oracle's double instantiation: no
// floating-point contraction, identical operation order.
#pragma clang fp contract(off)


struct IBox {
  double lo[3], hi[3];  // t, u, v
};

// std::min / std::max semantics (first argument wins ties and NaN).
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

__device__ __forceinline__ double linf3(const double a[3], const double b[3]) {
  double m = 0.0;
  for (int i = 0; i < 3; ++i) m = dmax(m, fabs(b[i] - a[i]));
  return m;
}

__device__ __forceinline__ double clamp_div(double a, double b) {
  if (b == 0.0) return __builtin_inf();
  return dmin(a / b, __builtin_inf());
}

__device__ void param_tolerance(bool vf, const double p0[4][3],
                                const double p1[4][3], double dist,
                                double tol[3]) {
  double q[2][4][3];
  for (int t = 0; t < 2; ++t) {
    const double(*p)[3] = t == 0 ? p0 : p1;
    for (int d = 0; d < 3; ++d) {
      if (vf) {
        q[t][0][d] = p[0][d] - p[1][d];
        q[t][1][d] = p[0][d] - p[3][d];
        q[t][2][d] = p[0][d] - (p[2][d] + p[3][d] - p[1][d]);
        q[t][3][d] = p[0][d] - p[2][d];
      } else {
        q[t][0][d] = p[0][d] - p[2][d];
        q[t][1][d] = p[0][d] - p[3][d];
        q[t][2][d] = p[1][d] - p[3][d];
        q[t][3][d] = p[1][d] - p[2][d];
      }
    }
  }
  auto m4 = [](const double* a1, const double* b1, const double* a2,
               const double* b2, const double* a3, const double* b3,
               const double* a4, const double* b4) {
    return dmax(dmax(linf3(a1, b1), linf3(a2, b2)),
                dmax(linf3(a3, b3), linf3(a4, b4)));
  };
  const double* p000 = q[0][0]; const double* p001 = q[0][1];
  const double* p011 = q[0][2]; const double* p010 = q[0][3];
  const double* p100 = q[1][0]; const double* p101 = q[1][1];
  const double* p111 = q[1][2]; const double* p110 = q[1][3];
  const double dl = 3 * m4(p000, p100, p001, p101, p011, p111, p010, p110);
  const double e0 = 3 * m4(p000, p010, p100, p110, p101, p111, p001, p011);
  const double e1 = 3 * m4(p000, p001, p100, p101, p110, p111, p010, p011);
  tol[0] = clamp_div(dist, dl);
  tol[1] = clamp_div(dist, e0);
  tol[2] = clamp_div(dist, e1);
}

// Range of one coordinate of the distance function over the unit box: the
// extremes of its values at the box corners (the unit-box evaluation of
// interval_root_finder.cpp), independent of the minimum separation.
__device__ void unit_range(bool vf, const double p0[4][3], const double p1[4][3],
                           int d, double* mn_out, double* mx_out) {
  double mn, mx;
  if (vf) {
    const double D[6] = {p0[0][d] - p0[1][d], p0[0][d] - p0[2][d],
                         p0[0][d] - p0[3][d], p1[0][d] - p1[1][d],
                         p1[0][d] - p1[2][d], p1[0][d] - p1[3][d]};
    mn = D[0];
    mx = D[0];
    for (int i = 1; i < 6; ++i) {
      if (D[i] < mn) mn = D[i];
      if (mx < D[i]) mx = D[i];
    }
  } else {
    const double D[8] = {p0[0][d] - p0[2][d], p0[0][d] - p0[3][d],
                         p0[1][d] - p0[2][d], p0[1][d] - p0[3][d],
                         p1[0][d] - p1[2][d], p1[0][d] - p1[3][d],
                         p1[1][d] - p1[2][d], p1[1][d] - p1[3][d]};
    mn = D[0];
    mx = D[0];
    for (int i = 1; i < 8; ++i) {
      if (D[i] < mn) mn = D[i];
      if (mx < D[i]) mx = D[i];
    }
  }
  *mn_out = mn;
  *mx_out = mx;
}

// Range [lo, hi] of one coordinate of the distance function over the
// (u, v) rectangle of a box at one time t (the box evaluation of
// interval_root_finder.cpp, one end of its t interval); on a VF query the
// rectangle is clipped to u + v <= 1.  c0..c5: that coordinate's affine
// coefficients.
__device__ void end_range(bool vf, double t, double ulo, double uhi, double vlo,
                          double vhi, const double c[6], double* lo_out,
                          double* hi_out) {
  const double q0 = (c[1] - c[0]) * t + c[0];
  const double q1 = (c[3] - c[2]) * t + c[2];
  const double q2 = (c[5] - c[4]) * t + c[4];
  double lo, hi;
  if (!vf || uhi + vhi <= 1.0) {
    const double a = q1 * ulo, bb = q1 * uhi, cc = q2 * vlo, dd = q2 * vhi;
    lo = q0 + dmin(a, bb) + dmin(cc, dd);
    hi = q0 + dmax(a, bb) + dmax(cc, dd);
  } else {
    lo = __builtin_inf();
    hi = -__builtin_inf();
    auto inc = [&](double u, double v) {
      const double val = q0 + q1 * u + q2 * v;
      lo = dmin(lo, val);
      hi = dmax(hi, val);
    };
    if (ulo + vlo <= 1.0) inc(ulo, vlo);
    if (ulo + vhi <= 1.0) inc(ulo, vhi);
    if (uhi + vlo <= 1.0) inc(uhi, vlo);
    if (uhi + vhi <= 1.0) inc(uhi, vhi);
    const double v_at_ulo = 1.0 - ulo;
    if (v_at_ulo >= vlo && v_at_ulo <= vhi) inc(ulo, v_at_ulo);
    const double v_at_uhi = 1.0 - uhi;
    if (v_at_uhi >= vlo && v_at_uhi <= vhi) inc(uhi, v_at_uhi);
    const double u_at_vlo = 1.0 - vlo;
    if (u_at_vlo >= ulo && u_at_vlo <= uhi) inc(u_at_vlo, vlo);
    const double u_at_vhi = 1.0 - vhi;
    if (u_at_vhi >= ulo && u_at_vhi <= uhi) inc(u_at_vhi, vhi);
  }
  *lo_out = lo;
  *hi_out = hi;
}

// Result of evaluating one box (a pure function of the box for a fixed query
// and minimum separation): bit 0 the box may hold a root, bit 1 it meets the
// stopping criterion (all widths within the tolerance, or inside the
// separation band in every coordinate).
constexpr int kBoxHit = 1;
constexpr int kBoxSmall = 2;

__device__ int unit_flags(const double umn[3], const double umx[3],
                          const double err[3], double ms, double co_tol) {
  bool in_eps_all = true;
  double width[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d) {
    const double em = err[d] + ms;
    width[d] = umx[d] - umn[d];
    if (umn[d] > em || umx[d] < -em) return 0;
    in_eps_all = in_eps_all && (umn[d] >= -em && umx[d] <= em);
  }
  const bool small = (width[0] <= co_tol && width[1] <= co_tol &&
                      width[2] <= co_tol) || in_eps_all;
  return small ? (kBoxHit | kBoxSmall) : kBoxHit;
}

__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = static_cast<unsigned>(
      __builtin_amdgcn_readlane(static_cast<int>(b), l));
  const unsigned hi = static_cast<unsigned>(
      __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}

// a, b or c for i = 0, 1, 2, by masks on the bit patterns (a select chain on
// a lane-dependent index is compiled into an indexed private-memory load).
__device__ __forceinline__ double pick3(int i, double a, double b, double c) {
  const unsigned long long ma = 0ull - static_cast<unsigned long long>(i == 0);
  const unsigned long long mb = 0ull - static_cast<unsigned long long>(i == 1);
  const unsigned long long mc = 0ull - static_cast<unsigned long long>(i == 2);
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(__double_as_longlong(a)) & ma) |
      (static_cast<unsigned long long>(__double_as_longlong(b)) & mb) |
      (static_cast<unsigned long long>(__double_as_longlong(c)) & mc)));
}

// Value of the neighbouring lane (lanes 2i and 2i + 1 swap; DPP quad_perm
// [1, 0, 3, 2]).  All lanes must be active.
__device__ __forceinline__ double swap_pair(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = static_cast<unsigned>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0xB1, 0xF, 0xF, false));
  const unsigned hi = static_cast<unsigned>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0xB1, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}

static_assert(kCcdStackCap == 64, "one pending box per lane of a wavefront");

enum : int { kRootFound = 1, kRootNone = 0, kRootPruned = 2 };

// Read by every lane at once; taken from the first lane so that the value
// (and the branch on it) is uniform over the wavefront.
__device__ __forceinline__ double load_cutoff(const unsigned long long* bits) {
  const unsigned long long v = __hip_atomic_load(
      const_cast<unsigned long long*>(bits), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_AGENT);
  const unsigned lo = static_cast<unsigned>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(v)));
  const unsigned hi = static_cast<unsigned>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32)));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(hi) << 32) | lo));
}

// interval_root_finder_bucket_DFS (interval_root_finder.cpp:391-510) run by
// one wavefront.  The pending list (at most kCcdStackCap = 64 boxes) is
// spread over the lanes: lane i holds one box, whether it is in use, its
// evaluation (computed when it is pushed; it depends on the box alone) and
// its rank in processing order.  The order is that of the sequential search:
// smallest t.lower first, ties to the most recently pushed box.  A pop is a
// ballot for rank 0 and a read of that lane; a push ranks the new box with
// one ballot.  The two halves of a split box are evaluated by twelve lanes
// at once, one per (half, coordinate, end of the t interval); a pair swap
// joins the two ends and three ballots give each half's flags.  Results,
// including where the list overflows (the current box's t.lower is
// returned: an earlier, conservative TOI) and the iteration limit, are those
// of the sequential search.  All lanes must be active.
// Lane-dependent updates are selects, not branches: a branch on the lane
// lets the optimiser route lanes around the wavefront operations on their
// own, which splits the wavefront.
// cutoff (optional): the bit pattern of a running minimum TOI.  Popped
// t.lower values never decrease (a half's t.lower is at least its parent's)
// and the search returns the t.lower of a popped box, so once the popped
// t.lower reaches the minimum the result cannot lower it: kRootPruned.  The
// minimum is re-read every 32 passes.
__device__ int wave_root_find(bool vf, const double umn[3], const double umx[3],
                              const double k[6][3], const double tol[3],
                              double co_tol, const double err[3], double ms,
                              int max_iter, const unsigned long long* cutoff,
                              double* t_lo) {
  const int lane = static_cast<int>(__lane_id());
  // This lane's part of a split evaluation: half h, coordinate d, end e
  // (lanes 12 and up repeat lanes 0..11 and are ignored).
  const int part = lane % 12;
  const int h = part / 6, d = (part % 6) / 2, e = part & 1;
  double kd[6];
  for (int i = 0; i < 6; ++i) kd[i] = pick3(d, k[i][0], k[i][1], k[i][2]);
  const double em = pick3(d, err[0], err[1], err[2]) + ms;
  const bool lead = lane < 12 && e == 0;
  bool valid = false;
  int rank = 0;
  int flags = 0;
  double key = 0.0;  // t.lower of this lane's box
  IBox box;
  for (int i = 0; i < 3; ++i) {
    box.lo[i] = 0.0;
    box.hi[i] = 1.0;
  }
  IBox cur = box;
  int cur_flags = __builtin_amdgcn_readfirstlane(unit_flags(umn, umx, err, ms, co_tol));
  double cut = __builtin_inf(), cut_next = __builtin_inf();
  if (cutoff != nullptr) {
    cut = load_cutoff(cutoff);
    cut_next = cut;
  }
  int iter = 0;
  for (;;) {
    ++iter;
    if (cutoff != nullptr) {
      if ((iter & 31) == 0) {
        cut = cut_next;
        cut_next = load_cutoff(cutoff);
      }
      if (cur.lo[0] >= cut) return kRootPruned;
    }
    if (cur_flags & kBoxHit) {
      if ((cur_flags & kBoxSmall) || (max_iter > 0 && iter > max_iter)) {
        *t_lo = cur.lo[0];
        return kRootFound;
      }
      // find_next_split: the largest width / tolerance among the widths
      // above their tolerance.
      int split = 0;
      double best = -__builtin_inf();
      for (int i = 0; i < 3; ++i) {
        const double w = cur.hi[i] - cur.lo[i];
        const double q = (w > tol[i]) ? w / tol[i] : -__builtin_inf();
        if (q > best) {
          best = q;
          split = i;
        }
      }
      double slo = cur.lo[0], shi = cur.hi[0];
      for (int i = 1; i < 3; ++i)
        if (split == i) {
          slo = cur.lo[i];
          shi = cur.hi[i];
        }
      const double mid = (shi + slo) / 2;
      if (mid < 1e-12) {
        *t_lo = cur.lo[0];
        return kRootFound;
      }
      IBox half[2] = {cur, cur};  // second (upper) half, first (lower) half
      for (int i = 0; i < 3; ++i)
        if (split == i) {
          half[0].lo[i] = mid;
          half[1].hi[i] = mid;
        }
      bool push_second = true, push_first = true;
      if (vf && split != 0) {
        const double other = (split == 1) ? cur.lo[2] : cur.lo[1];
        push_second = mid + other <= 1.0;
        push_first = slo + other <= 1.0;
      }
      const unsigned long long used = __ballot(valid);
      const int need = static_cast<int>(push_second) + static_cast<int>(push_first);
      if (__popcll(used) + need > kCcdStackCap) {
        *t_lo = cur.lo[0];
        return kRootFound;
      }
      // Evaluate both halves: lane (h, d, e) takes coordinate d at one end
      // of half h's t interval; the even lane of each pair joins the ends.
      int f[2];
      {
        double blo[3], bhi[3];
        for (int i = 0; i < 3; ++i) {
          blo[i] = h == 0 ? half[0].lo[i] : half[1].lo[i];
          bhi[i] = h == 0 ? half[0].hi[i] : half[1].hi[i];
        }
        double lo, hi;
        end_range(vf, e == 0 ? blo[0] : bhi[0], blo[1], bhi[1], blo[2], bhi[2],
                  kd, &lo, &hi);
        const double lo_o = swap_pair(lo), hi_o = swap_pair(hi);
        const double mn = dmin(lo, lo_o), mx = dmax(hi, hi_o);
        const bool ok = !(mn > em || mx < -em);
        const bool in_eps = mn >= -em && mx <= em;
        const bool thin = (mx - mn) <= co_tol;
        const unsigned long long okm = __ballot(lead && ok);
        const unsigned long long epsm = __ballot(lead && in_eps);
        const unsigned long long thinm = __ballot(lead && thin);
        for (int c = 0; c < 2; ++c) {
          const unsigned long long m = 0x15ull << (6 * c);
          const bool hit = (okm & m) == m;
          const bool small = (thinm & m) == m || (epsm & m) == m;
          f[c] = hit ? (small ? (kBoxHit | kBoxSmall) : kBoxHit) : 0;
        }
      }
      unsigned long long free_lanes = ~used;
      // Push the second half, then the first (the most recent).
      for (int c = 0; c < 2; ++c) {
        if (!(c == 0 ? push_second : push_first)) continue;
        const double nkey = half[c].lo[0];
        const int l = __builtin_ctzll(free_lanes);
        free_lanes &= free_lanes - 1;
        const bool below = key < nkey;
        const int before = __popcll(__ballot(valid && below));
        const bool me = lane == l;
        rank = me ? before : rank + ((valid && !below) ? 1 : 0);
        valid = valid || me;
        flags = me ? f[c] : flags;
        key = me ? nkey : key;
        for (int i = 0; i < 3; ++i) {
          box.lo[i] = me ? half[c].lo[i] : box.lo[i];
          box.hi[i] = me ? half[c].hi[i] : box.hi[i];
        }
      }
    }
    const unsigned long long head = __ballot(valid && rank == 0);
    if (head == 0) return kRootNone;
    const int l = __builtin_ctzll(head);
    for (int i = 0; i < 3; ++i) {
      cur.lo[i] = read_lane(box.lo[i], l);
      cur.hi[i] = read_lane(box.hi[i], l);
    }
    cur_flags = __builtin_amdgcn_readlane(flags, l);
    rank -= (valid && lane != l) ? 1 : 0;
    valid = valid && lane != l;
  }
}

__device__ void coefficients(bool vf, const double p0[4][3],
                             const double p1[4][3], double k[6][3]) {
  for (int d = 0; d < 3; ++d) {
    if (vf) {
      k[0][d] = p0[0][d] - p0[1][d];
      k[1][d] = p1[0][d] - p1[1][d];
      k[2][d] = p0[1][d] - p0[2][d];
      k[3][d] = p1[1][d] - p1[2][d];
      k[4][d] = p0[1][d] - p0[3][d];
      k[5][d] = p1[1][d] - p1[3][d];
    } else {
      k[0][d] = p0[0][d] - p0[2][d];
      k[1][d] = p1[0][d] - p1[2][d];
      k[2][d] = p0[1][d] - p0[0][d];
      k[3][d] = p1[1][d] - p1[0][d];
      k[4][d] = p0[2][d] - p0[3][d];
      k[5][d] = p1[2][d] - p1[3][d];
    }
  }
}

__device__ void trajectory(const CollVerts& v, int gv, const CcdParams& cp,
                           double p0[3], double p1[3]) {
  for (int d = 0; d < 3; ++d) {
    const double a = v.p0[3 * gv + d], b = v.p1[3 * gv + d];
    if (gv < v.nvc) {
      p0[d] = a;
      p1[d] = b;
    } else {
      const double delta = b - a;
      p0[d] = a + cp.time_start * delta;
      p1[d] = p0[d] + cp.interval * delta;
    }
  }
}

// Candidate pairs of a step: query q < npt is a point-triangle pair, the
// rest edge-edge pairs (edges in canonical order by vertex key).
struct PairQueries {
  CollVerts v;
  PrimTables pt;
  const int2* ptp;
  const int* ptc;
  int pt_cap;
  const int2* eep;
  const int* eec;
  int ee_cap;
  CcdParams cp;
  __device__ int count() const { return min(*ptc, pt_cap) + min(*eec, ee_cap); }
  __device__ double ms() const { return cp.ms; }
  __device__ void load(int q, bool* vf, double p0[4][3], double p1[4][3],
                       double err[3]) const {
    const int npt = min(*ptc, pt_cap);
    int gv[4];
    if (q < npt) {
      const int2 pr = ptp[q];
      *vf = true;
      gv[0] = pt.point[pr.y];
      gv[1] = pt.tri[3 * pr.x];
      gv[2] = pt.tri[3 * pr.x + 1];
      gv[3] = pt.tri[3 * pr.x + 2];
    } else {
      const int2 pr = eep[q - npt];
      *vf = false;
      int a0 = pt.edge[2 * pr.x], a1 = pt.edge[2 * pr.x + 1];
      int b0 = pt.edge[2 * pr.y], b1 = pt.edge[2 * pr.y + 1];
      const int ka0 = min(a0, a1), ka1 = max(a0, a1);
      const int kb0 = min(b0, b1), kb1 = max(b0, b1);
      if (kb0 < ka0 || (kb0 == ka0 && kb1 < ka1)) {
        int t0 = a0, t1 = a1;
        a0 = b0; a1 = b1;
        b0 = t0; b1 = t1;
      }
      gv[0] = a0; gv[1] = a1; gv[2] = b0; gv[3] = b1;
    }
    for (int i = 0; i < 4; ++i) trajectory(v, gv[i], cp, p0[i], p1[i]);
    const double* e = *vf ? cp.err_vf : cp.err_ee;
    for (int d = 0; d < 3; ++d) err[d] = e[d];
  }
};

// Explicit queries: 12 doubles per end position set, 3 filter values each.
struct ExplicitQueries {
  int n;
  const int* is_vf;
  const double* p0;
  const double* p1;
  const double* err;
  double sep;
  __device__ int count() const { return n; }
  __device__ double ms() const { return sep; }
  __device__ void load(int q, bool* vf, double a[4][3], double b[4][3],
                       double e[3]) const {
    *vf = is_vf[q] != 0;
    for (int i = 0; i < 12; ++i) {
      a[i / 3][i % 3] = p0[12 * static_cast<size_t>(q) + i];
      b[i / 3][i % 3] = p1[12 * static_cast<size_t>(q) + i];
    }
    for (int d = 0; d < 3; ++d) e[d] = err[3 * static_cast<size_t>(q) + d];
  }
};

// Outputs are written by every lane of the wavefront with the same values
// (zero() is called by lane 0 only).
// Earliest TOI over the candidates (toi == 0 is not a collision).  Queries
// that cannot lower it are cut short.
struct MinToiOut {
  unsigned long long* toi_bits;
  int* zero_count;  // may be null
  __device__ const unsigned long long* cutoff() const { return toi_bits; }
  __device__ void put(int, bool hit, double toi) const {
    if (!hit || toi == 0.0) return;
    atomicMin(toi_bits, static_cast<unsigned long long>(__double_as_longlong(toi)));
  }
  __device__ void zero() const {
    if (zero_count != nullptr) atomicAdd(zero_count, 1);
  }
};

struct EachOut {
  int* hit;
  double* toi;
  __device__ const unsigned long long* cutoff() const { return nullptr; }
  __device__ void put(int q, bool h, double t) const {
    hit[q] = h ? 1 : 0;
    toi[q] = h ? t : 0.0;
  }
  __device__ void zero() const {}
};

// CCD, one query per wavefront at a time: wavefront w of W takes queries
// w, w + W, ... (the grid holds as many wavefronts as can be resident; a
// shared query counter serialised every wavefront on one atomic).  Per query
// (ccd.cpp:150-180, narrowphase.cu:68): the root search at ms; a root at
// t = 0 is searched again at ms / 2, where no root means a collision at t = 1
// and a root at t = 0 means none.  With SILK_HIP_CCD_ZERO_CONTACT a root at
// t = 0 at ms / 2 is searched a third time at zero separation, with the
// same meaning of its outcomes.
template <class In, class Out>
__global__ void __launch_bounds__(256)
    k_ccd(In in, Out out, double tolerance, int max_iter, int zero_mode) {
  const int n = in.count();
  const int lane = static_cast<int>(__lane_id());
  const int nwave = static_cast<int>(gridDim.x * blockDim.x) / dev::kWave;
  const int wave0 = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / dev::kWave);
  for (int q = __builtin_amdgcn_readfirstlane(wave0); q < n; q += nwave) {
    bool vf = false;
    double p0[4][3], p1[4][3], err[3];
    in.load(q, &vf, p0, p1, err);
    vf = __builtin_amdgcn_readfirstlane(static_cast<int>(vf)) != 0;
    double tol[3], k[6][3], umn[3], umx[3];
    param_tolerance(vf, p0, p1, tolerance, tol);
    coefficients(vf, p0, p1, k);
    for (int d = 0; d < 3; ++d) unit_range(vf, p0, p1, d, &umn[d], &umx[d]);
    const double ms = in.ms();
    const unsigned long long* cut = out.cutoff();
    bool hit = false, pruned = false, zero = false;
    double toi = 0.0, t = 0.0;
    const int r1 = wave_root_find(vf, umn, umx, k, tol, tolerance, err, ms,
                                  max_iter, cut, &t);
    if (r1 == kRootPruned) {
      pruned = true;
    } else if (r1 == kRootFound) {
      if (t > 0.0) {
        hit = true;
        toi = t;
      } else {
        // Searches again at the smaller separations while the root stays at
        // t = 0 (the loop bound and separation are wavefront-uniform).
        const int passes = zero_mode == 1 ? 2 : 1;
        zero = true;
        for (int pass = 0; pass < passes && zero; ++pass) {
          double t2 = 0.0;
          const double sep = pass == 0 ? ms * 0.5 : 0.0;
          const int r2 = wave_root_find(vf, umn, umx, k, tol, tolerance, err, sep,
                                        max_iter, cut, &t2);
          if (r2 == kRootPruned) {
            pruned = true;
            zero = false;
          } else if (r2 == kRootNone) {
            hit = true;
            toi = 1.0;
            zero = false;
          } else if (t2 > 0.0) {
            hit = true;
            toi = t2;
            zero = false;
          }
        }
      }
    }
    if (!pruned) out.put(q, hit, toi);
    if (zero && lane == 0) out.zero();
  }
}

#pragma clang fp contract(fast)

__global__ void __launch_bounds__(1024) k_ccd_filter(CollVerts v, double* err6) {
  __shared__ float red[3][1024];
  float m[3] = {0.0f, 0.0f, 0.0f};
  for (int i = threadIdx.x; i < v.nvt; i += blockDim.x)
    for (int d = 0; d < 3; ++d)
      m[d] = fmaxf(m[d], fmaxf(fabsf(v.p0[3 * i + d]), fabsf(v.p1[3 * i + d])));
  for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = m[d];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int d = 0; d < 3; ++d)
        red[d][threadIdx.x] = fmaxf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double vf = static_cast<double>(4.053116e-06f);
    const double ee = static_cast<double>(3.814698e-06f);
    for (int d = 0; d < 3; ++d) {
      const double g = fmax(static_cast<double>(red[d][0]), 1.0);
      err6[d] = vf * g;
      err6[3 + d] = ee * g;
    }
  }
}

// ---------------------------------------------------------------- contacts
constexpr int kContactThreads = 1024;
constexpr size_t kContactLdsBytes = 150 * 1024;

// Contact of round r for this thread (contacts are stored colour by colour;
// rounds run through the colours once for the warm start and then once per
// sweep).  Returns false when the thread has none in that round.
__device__ __forceinline__ bool round_contact(const ContactSet& cs, int r,
                                              int e_extra, ContactDev* out) {
  const int col = r % cs.ncolor;
  const int e = cs.color_off[col] + static_cast<int>(threadIdx.x) + e_extra;
  if (e >= cs.color_off[col + 1]) return false;
  *out = cs.c[e];
  return true;
}

// One Gauss-Seidel projection of contact c (contact_constraints.cu:245-297)
// against positions P; corr is the contact's correction, updated in place.
// The four positions of a contact, read unconditionally (an absent vertex
// reads vertex 0 and is not used), so that the reads go out together.
__device__ __forceinline__ void load_states(const ContactDev& c, const float* P,
                                            dev::f3* pv) {
#pragma unroll
  for (int s = 0; s < 4; ++s) pv[s] = dev::ld3(P, max(c.state[s], 0));
}

// Projection of the tangential part ut onto the friction disc of radius
// rad >= 0 (contact_constraints.cu:391-411): zero when |ut| <= rad, else
// ut scaled by 1 - rad/|ut|.  The test is made on |ut|^2 and the scale uses
// one reciprocal square root, which keeps the IEEE square-root and divide
// sequences off the colour round's dependent chain.
__device__ __forceinline__ dev::f3 friction_cone(const dev::f3& ut, float rad) {
  const float t2 = dev::dot3(ut, ut);
  if (t2 <= rad * rad) return dev::f3{0.0f, 0.0f, 0.0f};
  return (1.0f - rad * rsqrtf(t2)) * ut;
}

// Coulomb projection of a contact velocity u with normal n and friction mu
// (solve_coulomb_contact, contact_constraints.cu:391-411): a separating u
// (u.n >= 0) is kept; otherwise the normal part is removed and the
// tangential part projected onto the friction disc of radius -mu u.n.
__device__ __forceinline__ dev::f3 coulomb(const dev::f3& u, const dev::f3& n, float mu) {
  const float un = dev::dot3(u, n);
  if (un >= 0.0f) return u;
  return friction_cone(u - un * n, -mu * un);
}

// Normal and friction coefficient given apart from the record (the register
// kernel keeps them in LDS).
__device__ __forceinline__ void gs_core(const ContactDev& c, float* P,
                                        dev::f3& corr, dev::f3 n, float mu) {
  using namespace dev;
  f3 pv[4];
  load_states(c, P, pv);
  f3 u = {c.affine[0], c.affine[1], c.affine[2]};
#pragma unroll
  for (int s = 0; s < 4; ++s) u = u + c.coef[s] * pv[s];  // coef 0 if static
  const f3 old = corr;
  const f3 unew = coulomb(u - old, n, mu);
  const f3 du = unew - u;
  corr = old + du;
  for (int s = 0; s < 4; ++s) {
    if (c.gamma[s] == 0.0f) continue;
    st3(P, c.state[s], ld3(P, c.state[s]) + c.gamma[s] * du);
  }
}

// gs_core for a contact whose state vertices are distinct: the positions
// read for u are reused for the update (nothing else writes them during the
// round: contacts of one colour share no vertex), which saves a second read.
__device__ __forceinline__ void gs_core_once(const ContactDev& c, float* P,
                                             dev::f3& corr, dev::f3 n, float mu) {
  using namespace dev;
  f3 pv[4];
  load_states(c, P, pv);
  f3 u = {c.affine[0], c.affine[1], c.affine[2]};
#pragma unroll
  for (int s = 0; s < 4; ++s) u = u + c.coef[s] * pv[s];  // coef 0 if static
  const f3 old = corr;
  const f3 unew = coulomb(u - old, n, mu);
  const f3 du = unew - u;
  corr = old + du;
  for (int s = 0; s < 4; ++s) {
    if (c.gamma[s] == 0.0f) continue;
    st3(P, c.state[s], pv[s] + c.gamma[s] * du);
  }
}

__device__ __forceinline__ void gs_core(const ContactDev& c, float* P,
                                        dev::f3& corr) {
  gs_core(c, P, corr, dev::f3{c.normal[0], c.normal[1], c.normal[2]}, c.mu);
}

__device__ __forceinline__ bool distinct_states(const ContactDev& c) {
  for (int a = 0; a < 4; ++a)
    for (int b = a + 1; b < 4; ++b)
      if (c.state[a] >= 0 && c.state[a] == c.state[b]) return false;
  return true;
}

// Corrections C in LDS or global memory, indexed by contact id.
__device__ __forceinline__ void gs_project(const ContactDev& c, float* P,
                                           float* C) {
  dev::f3 corr = dev::ld3(C, c.id);
  gs_core(c, P, corr);
  dev::st3(C, c.id, corr);
}

__device__ __forceinline__ void warm_core(const ContactDev& c, float* P,
                                          const dev::f3& corr) {
  using namespace dev;
  for (int s = 0; s < 4; ++s) {
    if (c.gamma[s] == 0.0f) continue;
    st3(P, c.state[s], ld3(P, c.state[s]) + c.gamma[s] * corr);
  }
}

// warm_core for a contact whose state vertices are distinct (reads first).
__device__ __forceinline__ void warm_core_once(const ContactDev& c, float* P,
                                               const dev::f3& corr) {
  using namespace dev;
  f3 pv[4];
  load_states(c, P, pv);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (c.gamma[s] == 0.0f) continue;
    st3(P, c.state[s], pv[s] + c.gamma[s] * corr);
  }
}

__device__ __forceinline__ void warm_start(const ContactDev& c, float* P,
                                           const float* C) {
  warm_core(c, P, dev::ld3(C, c.id));
}

template <bool kLds>
__global__ void __launch_bounds__(kContactThreads)
    k_contact_update(ContactSet cs, const float* __restrict__ x, int nact) {
  using namespace dev;
  extern __shared__ float sh[];
  if (cs.skip != nullptr && *cs.skip != 0) return;
  float* P = kLds ? sh : cs.pc;
  float* C = kLds ? sh + 3 * static_cast<size_t>(nact) : cs.corr;
  // q = x - eta / w for active contact vertices, x for the others
  // (contact_constraints.cu:203-214); the previous p stays in cs.p for the
  // dual residual.
  for (int l = threadIdx.x; l < nact; l += blockDim.x) {
    const int i = cs.act[l];
    const float w = cs.w[i];
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      P[3 * l + d] = w > 0.0f ? x[k] - cs.eta[k] / w : x[k];
    }
  }
  if (kLds)
    for (int k = threadIdx.x; k < 3 * cs.n; k += blockDim.x) C[k] = cs.corr[k];
  __syncthreads();
  // Warm start (one pass over the colours), then the sweeps; the next
  // round's contact is fetched before the barrier of the current one.
  const int rounds = cs.ncolor * (1 + cs.sweeps);
  ContactDev cur, nxt;
  bool have = rounds > 0 && round_contact(cs, 0, 0, &cur);
  for (int r = 0; r < rounds; ++r) {
    const bool have_next = r + 1 < rounds && round_contact(cs, r + 1, 0, &nxt);
    const bool warm = r < cs.ncolor;
    if (have) {
      if (warm) warm_start(cur, P, C);
      else gs_project(cur, P, C);
    }
    // Colours larger than the workgroup: the remaining contacts of the
    // round, without prefetch.
    const int col = r % cs.ncolor;
    for (int extra = kContactThreads;
         cs.color_off[col] + extra < cs.color_off[col + 1]; extra += kContactThreads) {
      ContactDev c;
      if (round_contact(cs, r, extra, &c)) {
        if (warm) warm_start(c, P, C);
        else gs_project(c, P, C);
      }
    }
    __syncthreads();
    cur = nxt;
    have = have_next;
  }
  // Multiplier update and dual residual terms (contact_constraints.cu:316-364).
  for (int l = threadIdx.x; l < nact; l += blockDim.x) {
    const int i = cs.act[l];
    const float w = cs.w[i];
    for (int d = 0; d < 3; ++d) {
      const int k = 3 * i + d;
      const float pn = P[3 * l + d];
      const float po = cs.p[k];
      cs.p[k] = pn;
      if (w > 0.0f) {
        cs.eta[k] += w * (pn - x[k]);
        cs.dual[k] = w * (pn - po);
      }
    }
  }
  if (kLds)
    for (int k = threadIdx.x; k < 3 * cs.n; k += blockDim.x) cs.corr[k] = C[k];
}

// Register-resident variant: contact e (colour order) and its correction
// live in thread e % kContactRegThreads for the whole update, so the rounds
// touch LDS only and their barriers wait on LDS traffic alone (a barrier
// also drains outstanding global loads, which made every round pay one).
// Contacts of one colour are independent, so a thread holding two of the
// same colour may project them in either order.
// With kNmLds each contact's normal and friction coefficient sit in LDS
// (4 floats per contact after P and the offsets, 16-byte aligned) rather than
// in registers: that frees 16 VGPRs per thread, which the 1024-thread launch
// would otherwise spill, and a spill reload sat on every round's chain.
constexpr int kContactRegThreads = kContactThreads;
constexpr int kContactSlots = 4;

__host__ __device__ constexpr size_t contact_nm_word(int nv, int ncolor) {
  return (3 * static_cast<size_t>(nv) + static_cast<size_t>(ncolor) + 1 + 3) & ~size_t{3};
}

template <bool kNmLds>
__global__ void __launch_bounds__(kContactRegThreads)
    k_contact_update_reg(ContactSet cs, const float* __restrict__ x, int nact) {
  using namespace dev;
  extern __shared__ float sh[];
  if (cs.skip != nullptr && *cs.skip != 0) return;
  // One group per workgroup when the contacts are stored by group: its
  // contacts, colour ranges and contact vertices; state indices are made
  // relative to the group's first contact vertex when the contacts load.
  int n = cs.n, cbase = 0, abase = 0;
  const int* color_off = cs.color_off;
  if (cs.ngroup > 1) {
    const int b = static_cast<int>(blockIdx.x), K = cs.ncolor;
    cbase = cs.group_off[b * K];
    n = cs.group_off[(b + 1) * K] - cbase;
    abase = cs.group_act[b];
    nact = cs.group_act[b + 1] - abase;
    color_off = cs.group_off + b * K;
  }
  const int* act = cs.act + abase;
  float* P = sh;
  int* off = reinterpret_cast<int*>(sh + 3 * static_cast<size_t>(nact));  // ncolor + 1
  float* NM = sh + contact_nm_word(nact, cs.ncolor);  // kNmLds: 4 per contact
  const int tid = static_cast<int>(threadIdx.x);
  const int n3 = 3 * nact;
  // Per coordinate of the contact vertices, with the loads of four steps in
  // flight at once.
  for (int k0 = tid; k0 < n3; k0 += 4 * kContactRegThreads) {
    float xv[4], ev[4], wv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kl = min(k0 + u * kContactRegThreads, n3 - 1);
      const int i = act[kl / 3];
      const int k = 3 * i + kl % 3;
      xv[u] = x[k];
      ev[u] = cs.eta[k];
      wv[u] = cs.w[i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kl = k0 + u * kContactRegThreads;
      if (kl < n3) P[kl] = wv[u] > 0.0f ? xv[u] - ev[u] / wv[u] : xv[u];
    }
  }
  for (int i = tid; i <= cs.ncolor; i += kContactRegThreads) off[i] = color_off[i] - cbase;
  __syncthreads();
  ContactDev c[kContactSlots];
  f3 corr[kContactSlots];
  int color[kContactSlots];
  unsigned once = 0;  // bit s: contact s has distinct state vertices
#pragma unroll
  for (int s = 0; s < kContactSlots; ++s) {
    const int e = s * kContactRegThreads + tid;
    color[s] = -1;
    corr[s] = f3{0.0f, 0.0f, 0.0f};
    if (e < n) {
      c[s] = cs.c[cbase + e];
      if (abase != 0)
        for (int k = 0; k < 4; ++k)
          if (c[s].state[k] >= 0) c[s].state[k] -= abase;
      if constexpr (kNmLds)
        *reinterpret_cast<float4*>(NM + 4 * e) =
            float4{c[s].normal[0], c[s].normal[1], c[s].normal[2], c[s].mu};
      if (distinct_states(c[s])) once |= 1u << s;
      corr[s] = ld3(cs.corr, c[s].id);
      int lo = 0, hi = cs.ncolor;  // off[lo] <= e < off[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid;
      }
      color[s] = lo;
    }
  }
  __syncthreads();
  // Warm start (one pass over the colours), then the sweeps.
  const int rounds = cs.ncolor * (1 + cs.sweeps);
  int col = 0;
  for (int r = 0; r < rounds; ++r) {
    const bool warm = r < cs.ncolor;
#pragma unroll
    for (int s = 0; s < kContactSlots; ++s) {
      if (color[s] != col) continue;
      if (warm) {
        if (once & (1u << s)) warm_core_once(c[s], P, corr[s]);
        else warm_core(c[s], P, corr[s]);
        continue;
      }
      f3 n;
      float mu;
      if constexpr (kNmLds) {
        const float4 q = *reinterpret_cast<const float4*>(
            NM + 4 * (s * kContactRegThreads + tid));
        n = f3{q.x, q.y, q.z};
        mu = q.w;
      } else {
        n = f3{c[s].normal[0], c[s].normal[1], c[s].normal[2]};
        mu = c[s].mu;
      }
      if (once & (1u << s)) gs_core_once(c[s], P, corr[s], n, mu);
      else gs_core(c[s], P, corr[s], n, mu);
    }
    __syncthreads();
    if (++col == cs.ncolor) col = 0;
  }
  // Multiplier update and dual residual terms (contact_constraints.cu:316-364),
  // per coordinate with four steps' loads in flight.
  for (int k0 = tid; k0 < n3; k0 += 4 * kContactRegThreads) {
    float wv[4], po[4], xv[4], ev[4];
    int kg[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kl = min(k0 + u * kContactRegThreads, n3 - 1);
      const int i = act[kl / 3];
      kg[u] = 3 * i + kl % 3;
      wv[u] = cs.w[i];
      po[u] = cs.p[kg[u]];
      xv[u] = x[kg[u]];
      ev[u] = cs.eta[kg[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kl = k0 + u * kContactRegThreads;
      if (kl >= n3) continue;
      const int k = kg[u];
      const float pn = P[kl];
      cs.p[k] = pn;
      if (wv[u] > 0.0f) {
        cs.eta[k] = ev[u] + wv[u] * (pn - xv[u]);
        cs.dual[k] = wv[u] * (pn - po[u]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kContactSlots; ++s)
    if (color[s] >= 0) st3(cs.corr, c[s].id, corr[s]);
}

// The contact kernels' Coulomb projection on explicit cases (3 floats per
// case for u, n and out).
__global__ void k_coulomb_each(int n, const float* __restrict__ u,
                               const float* __restrict__ nrm,
                               const float* __restrict__ mu, float* __restrict__ out) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  dev::st3(out, i, coulomb(dev::ld3(u, i), dev::ld3(nrm, i), mu[i]));
}

}  // namespace

void launch_coulomb_each(hipStream_t s, int n, const float* u, const float* nrm,
                         const float* mu, float* out) {
  if (n <= 0) return;
  k_coulomb_each<<<(n + 255) / 256, 256, 0, s>>>(n, u, nrm, mu, out);
}

void launch_refit(hipStream_t s, const CollVerts& v, const TreeDev* trees_dev,
                  int ntrees, const int* leaf_block_off, int nblocks) {
  if (ntrees == 0 || nblocks == 0) return;
  k_refit_leaves<<<nblocks, kRefitLeaves, 0, s>>>(v, trees_dev, leaf_block_off, ntrees);
  k_refit_top<<<ntrees, kRefitThreads, 0, s>>>(trees_dev);
}

void launch_traverse_jobs(hipStream_t s, const CollVerts& v, const TraverseJob* jobs,
                          int njobs, int nitems, PairOut pt, PairOut ee) {
  if (njobs <= 0 || nitems <= 0) return;
  k_broadphase<<<(nitems + kTraverseThreads - 1) / kTraverseThreads,
                 kTraverseThreads, 0, s>>>(v, jobs, njobs, pt, ee);
}

void launch_dcd(hipStream_t s, const CollVerts& v, const PrimTables& pt,
                const int2* pt_pairs, const int* pt_count, int pt_cap,
                const int2* ee_pairs, const int* ee_count, int ee_cap,
                float time, float dhat, DcdRec* out, int* out_count,
                int out_cap) {
  k_dcd<<<1024, 256, 0, s>>>(v, pt, pt_pairs, pt_count, pt_cap, ee_pairs,
                             ee_count, ee_cap, time, dhat, out, out_count,
                             out_cap);
}

void launch_ccd_filter(hipStream_t s, const CollVerts& v, double* err6) {
  k_ccd_filter<<<1, 1024, 0, s>>>(v, err6);
}

// Workgroups of a CCD launch: as many as can be resident at once.
template <class In, class Out>
int ccd_blocks() {
  static const int blocks = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_ccd<In, Out>, 256, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        nb <= 0 || cus <= 0) {
      (void)hipGetLastError();
      return kCcdBlocks;
    }
    return nb * cus;
  }();
  return blocks;
}

void launch_ccd(hipStream_t s, const CollVerts& v, const PrimTables& pt,
                const int2* pt_pairs, const int* pt_count, int pt_cap,
                const int2* ee_pairs, const int* ee_count, int ee_cap,
                const CcdParams& cp, unsigned long long* toi_bits) {
  PairQueries in{v, pt, pt_pairs, pt_count, pt_cap, ee_pairs, ee_count, ee_cap, cp};
  k_ccd<<<ccd_blocks<PairQueries, MinToiOut>(), 256, 0, s>>>(
      in, MinToiOut{toi_bits, cp.zero_count}, cp.tolerance, cp.max_iter, cp.zero_mode);
}

void launch_ccd_each(hipStream_t s, int n, const int* is_vf, const double* p0,
                     const double* p1, const double* err, double ms,
                     double tolerance, int max_iter, int* hit, double* toi,
                     int zero_mode) {
  if (n <= 0) return;
  ExplicitQueries in{n, is_vf, p0, p1, err, ms};
  const int blocks = std::min(ccd_blocks<ExplicitQueries, EachOut>(), (n + 3) / 4);
  k_ccd<<<blocks, 256, 0, s>>>(in, EachOut{hit, toi}, tolerance, max_iter, zero_mode);
}

void launch_ccd_min(hipStream_t s, int n, const int* is_vf, const double* p0,
                    const double* p1, const double* err, double ms,
                    double tolerance, int max_iter,
                    unsigned long long* toi_bits, int zero_mode) {
  if (n <= 0) return;
  ExplicitQueries in{n, is_vf, p0, p1, err, ms};
  const int blocks = std::min(ccd_blocks<ExplicitQueries, MinToiOut>(), (n + 3) / 4);
  k_ccd<<<blocks, 256, 0, s>>>(in, MinToiOut{toi_bits, nullptr}, tolerance, max_iter,
                               zero_mode);
}

namespace {

size_t contact_p_bytes(int nact, int ncolor) {
  return 12 * static_cast<size_t>(nact) + 4 * (static_cast<size_t>(ncolor) + 1);
}
size_t contact_nm_bytes(int nact, int n, int ncolor) {
  return 4 * contact_nm_word(nact, ncolor) + 16 * static_cast<size_t>(n);
}
size_t contact_lds_bytes(int nact, int n) {
  return 12 * (static_cast<size_t>(nact) + static_cast<size_t>(n));
}
bool contact_variant_fits(int v, int nact, int n, int ncolor) {
  switch (v) {
    case kContactRegNm:
      return n <= kContactSlots * kContactRegThreads &&
             contact_nm_bytes(nact, n, ncolor) <= kContactLdsBytes;
    case kContactReg:
      return n <= kContactSlots * kContactRegThreads &&
             contact_p_bytes(nact, ncolor) <= kContactLdsBytes;
    case kContactLds:
      return contact_lds_bytes(nact, n) <= kContactLdsBytes;
    default:
      return true;
  }
}

}  // namespace

bool contact_group_fits(int nact, int n, int ncolor) {
  return contact_variant_fits(kContactRegNm, nact, n, ncolor);
}

int contact_update_variant(int nact, int n, int ncolor) {
  for (int v = kContactRegNm; v < kContactGlobal; ++v)
    if (contact_variant_fits(v, nact, n, ncolor)) return v;
  return kContactGlobal;
}

void launch_contact_update(hipStream_t s, const ContactSet& cs,
                           const float* x) {
  if (cs.n == 0 || cs.ncolor == 0) return;
  int variant = contact_update_variant(cs.nact, cs.n, cs.ncolor);
  // A forced variant is taken only where it fits (the global one always does).
  if (cs.force_variant >= 0 &&
      contact_variant_fits(cs.force_variant, cs.nact, cs.n, cs.ncolor))
    variant = cs.force_variant;
  const size_t p_bytes = contact_p_bytes(cs.nact, cs.ncolor);
  const size_t nm_bytes = contact_nm_bytes(cs.nact, cs.n, cs.ncolor);
  if (cs.ngroup > 1) {
    // One workgroup per independent group; the caller groups the contacts
    // only when every group fits the register variant.
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_contact_update_reg<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kContactLdsBytes));
      attr = true;
    }
    const size_t bytes = contact_nm_bytes(cs.group_nact_max, cs.group_n_max, cs.ncolor);
    k_contact_update_reg<true><<<cs.ngroup, kContactRegThreads, bytes, s>>>(cs, x, cs.nact);
    return;
  }
  if (variant == kContactRegNm) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_contact_update_reg<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kContactLdsBytes));
      attr = true;
    }
    k_contact_update_reg<true><<<1, kContactRegThreads, nm_bytes, s>>>(cs, x, cs.nact);
  } else if (variant == kContactReg) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_contact_update_reg<false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kContactLdsBytes));
      attr = true;
    }
    k_contact_update_reg<false><<<1, kContactRegThreads, p_bytes, s>>>(cs, x, cs.nact);
  } else if (variant == kContactLds) {
    const size_t shmem = contact_lds_bytes(cs.nact, cs.n);
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_contact_update<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(kContactLdsBytes));
      attr = true;
    }
    k_contact_update<true><<<1, kContactThreads, shmem, s>>>(cs, x, cs.nact);
  } else {
    k_contact_update<false><<<1, kContactThreads, 0, s>>>(cs, x, cs.nact);
  }
}

}  // namespace silk::amd
