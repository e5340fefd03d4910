// The reference CPU solver restated as a timing baseline (TEST AND BASELINE
// INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and tests/,
// never by the product library).
//
// Follows silk/src/backend/cpu/solver/pipeline.cpp:45-168 (outer CCD loop,
// inner PD loop, barrier reflection targets, rollback to 0.8 TOI),
// cloth_solver_context.cpp:21-73 (H = M/h^2 + k JWJ + kb CWC + pin penalty,
// factored once per dt), cloth_solver_utils.cpp:163-294 (outer right-hand
// side, per-face SVD projection clamped to [0.9, 1.1], global solve) and
// backend/cpu/collision/pipeline.cpp:110-245 with narrowphase.cpp:120-263
// (DCD with reflection corrections, FP32 interval-root-finder CCD).  FP32
// state as in the reference.
//
// Own choices where the reference calls a library: the Cholesky factor is a
// project-written envelope (profile) factorization of the vertex-scalar
// matrix in reverse Cuthill-McKee order (H is a vertex-scalar matrix times
// I3, pin and barrier terms included, so one scalar factor serves the three
// coordinates; the reference factors the 3N matrix with CHOLMOD and updates
// it with cholmod_updown when barriers appear -- here the scalar matrix with
// the barrier diagonal is refactored); the broadphase is sort-and-sweep on
// padded swept boxes (the reference uses a KD-tree) with the same filters;
// threads are std::thread workers over faces and candidates (the reference
// uses TBB).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <queue>
#include <thread>
#include <vector>

#include "oracle.hpp"

namespace oracle {
namespace cpu {

namespace {

constexpr float kPinPenalty = 1e4f;        // cloth_solver_context.cpp:17
constexpr float kBarrierStiffness = 1e4f;  // collision/pipeline.hpp:29
constexpr float kDcdEps = 1e-6f;           // narrowphase.cpp:9

// Run f(begin, end) over [0, n) on `threads` workers (the calling thread
// takes the first chunk).
void parallel_for(int threads, int64_t n, const std::function<void(int64_t, int64_t, int)>& f) {
  if (n <= 0) return;
  const int t = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(threads, n)));
  if (t == 1) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (n + t - 1) / t;
  for (int k = 1; k < t; ++k) {
    const int64_t b = k * chunk, e = std::min(n, b + chunk);
    if (b < e) pool.emplace_back(f, b, e, k);
  }
  f(0, std::min(n, chunk), 0);
  for (auto& th : pool) th.join();
}

// Envelope Cholesky of a symmetric positive definite sparse matrix, rows in
// a given order.  Row i of L is stored densely from its first nonzero column.
struct Envelope {
  int n = 0;
  std::vector<int> first;        // first column of row i (permuted numbering)
  std::vector<int64_t> start;    // offset of row i's [first, i] entries
  std::vector<double> L;
  std::vector<int> perm, iperm;  // perm[new] = old, iperm[old] = new

  // Symbolic: the envelope of A (upper pattern by row) in the new order.
  void analyse(const std::vector<std::vector<int>>& adj, const std::vector<int>& p) {
    n = static_cast<int>(adj.size());
    perm = p;
    iperm.assign(n, 0);
    for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
    first.assign(n, 0);
    start.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      int f = i;
      for (int j : adj[perm[i]]) f = std::min(f, iperm[j]);
      first[i] = f;
      start[i + 1] = start[i] + (i - f + 1);
    }
    L.assign(static_cast<size_t>(start[n]), 0.0);
  }

  // Numeric: rows of A given as (col, val) lists in the original numbering
  // plus an extra diagonal.  False when a pivot is not positive.
  bool factor(const std::vector<std::vector<std::pair<int, double>>>& rows,
              const std::vector<double>& diag_add) {
    std::fill(L.begin(), L.end(), 0.0);
    for (int i = 0; i < n; ++i) {
      double* li = &L[static_cast<size_t>(start[i])];
      const int fi = first[i];
      for (const auto& [c, v] : rows[perm[i]]) {
        const int j = iperm[c];
        if (j <= i) li[j - fi] += v;
      }
      li[i - fi] += diag_add[perm[i]];
      for (int j = fi; j <= i; ++j) {
        const double* lj = &L[static_cast<size_t>(start[j])];
        const int fj = first[j];
        const int k0 = std::max(fi, fj);
        double s = li[j - fi];
        for (int k = k0; k < j; ++k) s -= li[k - fi] * lj[k - fj];
        if (j < i) {
          li[j - fi] = s / lj[j - fj];
        } else {
          if (!(s > 0.0)) return false;
          li[i - fi] = std::sqrt(s);
        }
      }
    }
    return true;
  }

  // Solve L L^T x = b for three right-hand sides (x, y, z interleaved per
  // vertex in the original numbering).
  void solve3(const std::vector<double>& b, std::vector<double>& x) const {
    std::vector<double> y(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
      const double* li = &L[static_cast<size_t>(start[i])];
      const int fi = first[i];
      double s0 = b[3 * perm[i]], s1 = b[3 * perm[i] + 1], s2 = b[3 * perm[i] + 2];
      for (int k = fi; k < i; ++k) {
        const double l = li[k - fi];
        s0 -= l * y[3 * k];
        s1 -= l * y[3 * k + 1];
        s2 -= l * y[3 * k + 2];
      }
      const double d = li[i - fi];
      y[3 * i] = s0 / d;
      y[3 * i + 1] = s1 / d;
      y[3 * i + 2] = s2 / d;
    }
    for (int i = n - 1; i >= 0; --i) {
      const double* li = &L[static_cast<size_t>(start[i])];
      const int fi = first[i];
      const double d = li[i - fi];
      y[3 * i] /= d;
      y[3 * i + 1] /= d;
      y[3 * i + 2] /= d;
      for (int k = fi; k < i; ++k) {
        const double l = li[k - fi];
        y[3 * k] -= l * y[3 * i];
        y[3 * k + 1] -= l * y[3 * i + 1];
        y[3 * k + 2] -= l * y[3 * i + 2];
      }
    }
    x.assign(3 * static_cast<size_t>(n), 0.0);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) x[3 * perm[i] + d] = y[3 * i + d];
  }
};

// Reverse Cuthill-McKee order of a graph, started in every component from a
// minimum-degree vertex.
std::vector<int> rcm_order(const std::vector<std::vector<int>>& adj) {
  const int n = static_cast<int>(adj.size());
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  std::vector<int> byd(n);
  for (int i = 0; i < n; ++i) byd[i] = i;
  std::stable_sort(byd.begin(), byd.end(),
                   [&](int a, int b) { return adj[a].size() < adj[b].size(); });
  for (int s : byd) {
    if (seen[s]) continue;
    std::queue<int> q;
    q.push(s);
    seen[s] = 1;
    while (!q.empty()) {
      const int v = q.front();
      q.pop();
      order.push_back(v);
      std::vector<int> nb;
      for (int u : adj[v])
        if (!seen[u]) {
          seen[u] = 1;
          nb.push_back(u);
        }
      std::sort(nb.begin(), nb.end(),
                [&](int a, int b) { return adj[a].size() < adj[b].size(); });
      for (int u : nb) q.push(u);
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

struct Collider {
  int type = 0;     // 0 point, 1 edge, 2 triangle
  int v[3] = {0, 0, 0};
  float inv_mass[3] = {0, 0, 0};
};

struct Object {
  bool dynamic = false;
  int nv = 0, off = 0;            // vertices; offset into the global position arrays
  std::vector<float> V0;          // rest positions
  std::vector<int> F, E;
  float padding = 0.0f;
  std::vector<Collider> points, edges, tris;
  // cloth only
  std::vector<char> pinned;
  std::vector<float> pin_pos;     // 3 per pinned vertex (caller order of pins)
  std::vector<int> pins;
  std::vector<float> mass;        // density * Voronoi mass
  std::vector<std::array<float, 54>> jop;
  std::vector<float> area;
  std::vector<std::vector<std::pair<int, double>>> H;  // scalar rows
  std::vector<std::vector<int>> adj;
  std::vector<double> c0;         // kb * CWC * V (3 per vertex)
  Envelope chol;
  bool plain = false;             // chol holds H without barrier terms
  std::vector<float> x, v;        // state (3 nv)
  // obstacle only
  std::vector<float> prev, curr;
};

struct Collision {
  int type = 0;          // 0 PT, 1 EE
  int gv[4];             // global vertex ids
  float inv_mass[4];
  float corr[4][3];      // correction per vertex (weight_i (dhat - d) n)
};

}  // namespace

struct Pipeline {
  int threads = 1;
  float dt = 1.0f / 60.0f, acc[3] = {0, 0, -9.81f};
  float k = 100.0f, kb = 1e-4f, density = 0.1f, damping = 0.01f;
  int max_outer = 100, max_inner = 10, ccd_max_iter = 4096;
  float ms_scale = 0.2f, dcd_scale = 1.0f, ccd_tol = 1e-6f;
  std::vector<Object> objs;  // cloths first (insertion order), then obstacles
  // stats of the last step
  int outer = 0, inner = 0, collisions = 0, ccd_candidates = 0;
  double min_toi = 1.0;
};

namespace {

// Colliders of an object (object_collider.cpp:28-146): points, unique edges
// and triangles with per-vertex inverse masses (0 for pinned and obstacles).
void make_colliders(Object& o) {
  std::vector<std::pair<int, int>> e;
  for (size_t f = 0; f < o.F.size() / 3; ++f)
    for (int q = 0; q < 3; ++q) {
      const int a = o.F[3 * f + q], b = o.F[3 * f + (q + 1) % 3];
      e.emplace_back(std::min(a, b), std::max(a, b));
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  double len = 0.0;
  o.E.clear();
  for (auto& [a, b] : e) {
    o.E.push_back(a);
    o.E.push_back(b);
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double t = o.V0[3 * a + d] - o.V0[3 * b + d];
      s += t * t;
    }
    len += std::sqrt(s);
  }
  o.padding = 0.05f * static_cast<float>(e.empty() ? 0.0 : len / e.size());
  auto im = [&](int i) -> float {
    if (!o.dynamic || o.pinned[i]) return 0.0f;
    return 1.0f / o.mass[i];
  };
  o.points.clear();
  o.edges.clear();
  o.tris.clear();
  for (int i = 0; i < o.nv; ++i) {
    Collider c;
    c.type = 0;
    c.v[0] = i;
    c.inv_mass[0] = im(i);
    o.points.push_back(c);
  }
  for (size_t i = 0; i < o.E.size() / 2; ++i) {
    Collider c;
    c.type = 1;
    c.v[0] = o.E[2 * i];
    c.v[1] = o.E[2 * i + 1];
    c.inv_mass[0] = im(c.v[0]);
    c.inv_mass[1] = im(c.v[1]);
    o.edges.push_back(c);
  }
  for (size_t f = 0; f < o.F.size() / 3; ++f) {
    Collider c;
    c.type = 2;
    for (int q = 0; q < 3; ++q) {
      c.v[q] = o.F[3 * f + q];
      c.inv_mass[q] = im(c.v[q]);
    }
    o.tris.push_back(c);
  }
}

// Positions of object o at the segment ends (3 per vertex, object-local).
struct Frame {
  std::vector<const float*> t0, t1;  // per object
};

Box collider_box(const Object& o, const Collider& c, const float* p0, const float* p1) {
  Box b;
  const int n = c.type + 1;
  for (int d = 0; d < 3; ++d) {
    float lo = std::numeric_limits<float>::max(), hi = -lo;
    for (int q = 0; q < n; ++q) {
      lo = std::min({lo, p0[3 * c.v[q] + d], p1[3 * c.v[q] + d]});
      hi = std::max({hi, p0[3 * c.v[q] + d], p1[3 * c.v[q] + d]});
    }
    b.mn[d] = lo - o.padding;
    b.mx[d] = hi + o.padding;
  }
  return b;
}

struct Candidate {
  int oa, ob;        // objects
  const Collider* a;
  const Collider* b;
};

// Candidate pairs (collision/pipeline.cpp:23-158): object pairs of one group
// with at least one dynamic object -- triangles of one against points of the
// other both ways and edges against edges -- then self pairs of every cloth
// (non-adjacent, not all pinned).
std::vector<Candidate> candidates(const Pipeline& P, const Frame& fr) {
  std::vector<Candidate> out;
  const int no = static_cast<int>(P.objs.size());
  std::vector<std::vector<Box>> pb(no), eb(no), tb(no);
  for (int i = 0; i < no; ++i) {
    const Object& o = P.objs[i];
    for (const auto& c : o.points) pb[i].push_back(collider_box(o, c, fr.t0[i], fr.t1[i]));
    for (const auto& c : o.edges) eb[i].push_back(collider_box(o, c, fr.t0[i], fr.t1[i]));
    for (const auto& c : o.tris) tb[i].push_back(collider_box(o, c, fr.t0[i], fr.t1[i]));
  }
  auto pinned_all = [](const Collider& a, const Collider& b) {
    float s = 0.0f;
    for (int q = 0; q <= a.type; ++q) s += a.inv_mass[q];
    for (int q = 0; q <= b.type; ++q) s += b.inv_mass[q];
    return s == 0.0f;
  };
  for (int i = 0; i < no; ++i)
    for (int j = i + 1; j < no; ++j) {
      if (!P.objs[i].dynamic && !P.objs[j].dynamic) continue;
      const Object& A = P.objs[i];
      const Object& B = P.objs[j];
      for (auto [t, p] : overlap_pairs(tb[i], pb[j], false))
        if (!pinned_all(A.tris[t], B.points[p]))
          out.push_back({i, j, &A.tris[t], &B.points[p]});
      for (auto [t, p] : overlap_pairs(tb[j], pb[i], false))
        if (!pinned_all(B.tris[t], A.points[p]))
          out.push_back({j, i, &B.tris[t], &A.points[p]});
      for (auto [a, b] : overlap_pairs(eb[i], eb[j], false))
        if (!pinned_all(A.edges[a], B.edges[b]))
          out.push_back({i, j, &A.edges[a], &B.edges[b]});
    }
  for (int i = 0; i < no; ++i) {
    const Object& A = P.objs[i];
    if (!A.dynamic) continue;
    for (auto [t, p] : overlap_pairs(tb[i], pb[i], false)) {
      const Collider& T = A.tris[t];
      const int pv = A.points[p].v[0];
      if (pv == T.v[0] || pv == T.v[1] || pv == T.v[2]) continue;
      if (!pinned_all(T, A.points[p])) out.push_back({i, i, &T, &A.points[p]});
    }
    for (auto [a, b] : overlap_pairs(eb[i], eb[i], true)) {
      const Collider& X = A.edges[a];
      const Collider& Y = A.edges[b];
      if (X.v[0] == Y.v[0] || X.v[0] == Y.v[1] || X.v[1] == Y.v[0] || X.v[1] == Y.v[1])
        continue;
      if (!pinned_all(X, Y)) out.push_back({i, i, &X, &Y});
    }
  }
  return out;
}

using F3 = std::array<float, 3>;
inline F3 sub(const F3& a, const F3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline float fdot(const F3& a, const F3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// exact_pt_uv in FP32 (cpu/collision/dcd.cpp:7-38).
bool pt_uv32(const F3& x0, const F3& x1, const F3& x2, const F3& x3, float* u, float* v) {
  const F3 x21 = sub(x2, x1), x31 = sub(x3, x1), x01 = sub(x0, x1);
  const float a = fdot(x21, x21), b = fdot(x21, x31), c = fdot(x31, x31);
  const float d = fdot(x21, x01), e = fdot(x31, x01);
  const float det = a * c - b * b;
  const float area_eps = kDcdEps * std::max(a, c);
  if (det < area_eps * area_eps) return false;
  const float b1 = (c * d - b * e) / det, b2 = (a * e - b * d) / det;
  if (b1 < 0.0f || b2 < 0.0f || b1 + b2 > 1.0f) return false;
  *u = b1;
  *v = b2;
  return true;
}

// DCD of one candidate at frame time `time` (narrowphase.cpp:124-218):
// dynamic colliders are read at t1 (refit to the rolled-back state),
// obstacles interpolated along the frame.
bool dcd_one(const Pipeline& P, const Frame& fr, const Candidate& c, float time, float dhat,
             Collision* out) {
  auto pos = [&](int obj, int vtx) -> F3 {
    const Object& o = P.objs[obj];
    const float* a = fr.t0[obj] + 3 * vtx;
    const float* b = fr.t1[obj] + 3 * vtx;
    if (o.dynamic) return {b[0], b[1], b[2]};
    return {a[0] + time * (b[0] - a[0]), a[1] + time * (b[1] - a[1]),
            a[2] + time * (b[2] - a[2])};
  };
  F3 x[4];
  float invm[4], param[4];
  int gv[4];
  const bool pt = c.a->type == 2;
  if (pt) {
    const Collider& T = *c.a;
    const Collider& Q = *c.b;
    x[0] = pos(c.ob, Q.v[0]);
    gv[0] = P.objs[c.ob].off + Q.v[0];
    invm[0] = Q.inv_mass[0];
    for (int q = 0; q < 3; ++q) {
      x[1 + q] = pos(c.oa, T.v[q]);
      gv[1 + q] = P.objs[c.oa].off + T.v[q];
      invm[1 + q] = T.inv_mass[q];
    }
    float u, v;
    if (!pt_uv32(x[0], x[1], x[2], x[3], &u, &v)) return false;
    param[0] = 1.0f;
    param[1] = 1.0f - u - v;
    param[2] = u;
    param[3] = v;
  } else {
    for (int q = 0; q < 2; ++q) {
      x[q] = pos(c.oa, c.a->v[q]);
      gv[q] = P.objs[c.oa].off + c.a->v[q];
      invm[q] = c.a->inv_mass[q];
      x[2 + q] = pos(c.ob, c.b->v[q]);
      gv[2 + q] = P.objs[c.ob].off + c.b->v[q];
      invm[2 + q] = c.b->inv_mass[q];
    }
    V3 X[4];
    for (int q = 0; q < 4; ++q) X[q] = {x[q][0], x[q][1], x[q][2]};
    double u, v;
    if (!ee_uv(X[0], X[1], X[2], X[3], kDcdEps, &u, &v)) return false;
    const float uf = static_cast<float>(u), vf = static_cast<float>(v);
    param[0] = 1.0f - uf;
    param[1] = uf;
    param[2] = 1.0f - vf;
    param[3] = vf;
  }
  F3 pa, pb;
  if (pt) {
    pa = x[0];
    for (int d = 0; d < 3; ++d)
      pb[d] = param[1] * x[1][d] + param[2] * x[2][d] + param[3] * x[3][d];
  } else {
    for (int d = 0; d < 3; ++d) {
      pa[d] = param[0] * x[0][d] + param[1] * x[1][d];
      pb[d] = param[2] * x[2][d] + param[3] * x[3][d];
    }
  }
  const F3 disp = sub(pb, pa);
  const float dist = std::sqrt(fdot(disp, disp));
  if (dist >= dhat || dist == 0.0f) return false;
  const F3 n = {disp[0] / dist, disp[1] / dist, disp[2] / dist};
  float den = 0.0f;
  for (int q = 0; q < 4; ++q) den += param[q] * param[q] * invm[q];
  if (den == 0.0f) return false;
  Collision r;
  r.type = pt ? 0 : 1;
  for (int q = 0; q < 4; ++q) {
    float wgt = param[q] * invm[q] / den;
    if (pt ? q == 0 : q < 2) wgt = -wgt;
    r.gv[q] = gv[q];
    r.inv_mass[q] = invm[q];
    for (int d = 0; d < 3; ++d) r.corr[q][d] = (dhat - dist) * n[d] * wgt;
  }
  *out = r;
  return true;
}

}  // namespace

// ------------------------------------------------------------------ setup

static void prepare_cloth(Pipeline& P, Object& o) {
  std::vector<double> Vd(o.V0.begin(), o.V0.end());
  const L2Cache l2 = build_l2(Vd, o.F);
  o.mass.resize(o.nv);
  for (int i = 0; i < o.nv; ++i) o.mass[i] = static_cast<float>(P.density * l2.mass[i]);
  o.area.resize(l2.area.size());
  o.jop.resize(l2.jop.size());
  for (size_t f = 0; f < l2.area.size(); ++f) {
    o.area[f] = static_cast<float>(l2.area[f]);
    for (int q = 0; q < 54; ++q) o.jop[f][q] = static_cast<float>(l2.jop[f][q]);
  }
  // Scalar H: M/h^2 + k sum_f area J^T J (the x block of kron(B, I3)) + kb CWC + pins.
  std::vector<std::map<int, double>> H(o.nv);
  const double h2 = static_cast<double>(P.dt) * P.dt;
  for (int i = 0; i < o.nv; ++i) H[i][i] += o.mass[i] / h2;
  for (size_t f = 0; f < l2.area.size(); ++f)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += l2.jop[f][9 * r + 3 * a] * l2.jop[f][9 * r + 3 * b];
        H[o.F[3 * f + a]][o.F[3 * f + b]] += P.k * l2.area[f] * s;
      }
  for (int i = 0; i < o.nv; ++i)
    for (const auto& [j, v] : l2.cwc[i]) H[i][j] += P.kb * v;
  for (int i : o.pins) H[i][i] += kPinPenalty;
  o.H.assign(o.nv, {});
  o.adj.assign(o.nv, {});
  for (int i = 0; i < o.nv; ++i)
    for (const auto& [j, v] : H[i]) {
      o.H[i].emplace_back(j, v);
      if (j != i) o.adj[i].push_back(j);
    }
  o.chol.analyse(o.adj, rcm_order(o.adj));
  o.c0.resize(3 * static_cast<size_t>(o.nv));
  for (size_t i = 0; i < o.c0.size(); ++i) o.c0[i] = P.kb * l2.c0[i];
}

// ------------------------------------------------------------------- step

static bool step(Pipeline& P) {
  const int no = static_cast<int>(P.objs.size());
  P.outer = P.inner = P.collisions = P.ccd_candidates = 0;
  P.min_toi = 1.0;
  // Global state: every cloth's vertices, velocities damped by (1 - damping)
  // (pipeline.cpp:186-224).
  std::vector<int> cl;
  int n3 = 0;
  for (int i = 0; i < no; ++i)
    if (P.objs[i].dynamic) {
      cl.push_back(i);
      P.objs[i].off = n3 / 3;
      n3 += 3 * P.objs[i].nv;
    }
  int off = n3 / 3;
  for (int i = 0; i < no; ++i)
    if (!P.objs[i].dynamic) {
      P.objs[i].off = off;
      off += P.objs[i].nv;
    }
  std::vector<float> curr(n3), vel(n3);
  for (int i : cl) {
    const Object& o = P.objs[i];
    for (int q = 0; q < 3 * o.nv; ++q) {
      curr[3 * o.off + q] = o.x[q];
      vel[3 * o.off + q] = (1.0f - P.damping) * o.v[q];
    }
  }
  // Scene bbox and scale (pipeline.cpp:229-237).
  float bmin[3] = {1e30f, 1e30f, 1e30f}, bmax[3] = {-1e30f, -1e30f, -1e30f};
  for (int q = 0; q < n3; ++q) {
    bmin[q % 3] = std::min(bmin[q % 3], curr[q]);
    bmax[q % 3] = std::max(bmax[q % 3], curr[q]);
  }
  float scale2 = 0.0f;
  for (int d = 0; d < 3; ++d) scale2 += (bmax[d] - bmin[d]) * (bmax[d] - bmin[d]);
  const float scene_scale = std::sqrt(scale2);
  // Invariant rhs: pin penalty targets + rest curvature (cloth_solver_utils.cpp:101-115).
  std::vector<double> init_rhs(n3, 0.0);
  for (int i : cl) {
    const Object& o = P.objs[i];
    for (int q = 0; q < 3 * o.nv; ++q) init_rhs[3 * o.off + q] = o.c0[q];
    for (size_t p = 0; p < o.pins.size(); ++p)
      for (int d = 0; d < 3; ++d)
        init_rhs[3 * (o.off + o.pins[p]) + d] = kPinPenalty * o.pin_pos[3 * p + d];
  }
  float pad = std::numeric_limits<float>::max();
  for (const auto& o : P.objs) pad = std::min(pad, o.padding);
  const float dhat = P.dcd_scale * pad, ms = P.ms_scale * pad;
  // Positions of the collider frame: dynamic objects from the state arrays,
  // obstacles from their frame (prev -> curr).
  std::vector<float> seg0(n3), seg1(n3);
  auto frame = [&](const std::vector<float>& a, const std::vector<float>& b) {
    Frame fr;
    fr.t0.resize(no);
    fr.t1.resize(no);
    seg0 = a;
    seg1 = b;
    for (int i = 0; i < no; ++i) {
      const Object& o = P.objs[i];
      if (o.dynamic) {
        fr.t0[i] = seg0.data() + 3 * o.off;
        fr.t1[i] = seg1.data() + 3 * o.off;
      } else {
        fr.t0[i] = o.prev.data();
        fr.t1[i] = o.curr.data();
      }
    }
    return fr;
  };
  auto find_collisions = [&](float time) {
    const Frame fr = frame(curr, curr);
    const std::vector<Candidate> cand = candidates(P, fr);
    std::vector<std::vector<Collision>> part(P.threads);
    parallel_for(P.threads, static_cast<int64_t>(cand.size()),
                 [&](int64_t b, int64_t e, int t) {
                   Collision c;
                   for (int64_t i = b; i < e; ++i)
                     if (dcd_one(P, fr, cand[i], time, dhat, &c)) part[t].push_back(c);
                 });
    std::vector<Collision> all;
    for (auto& v : part) all.insert(all.end(), v.begin(), v.end());
    return all;
  };
  std::vector<Collision> coll = find_collisions(0.0f);
  float remaining = 1.0f;
  const float h = P.dt;
  for (int it = 0; it < P.max_outer; ++it) {
    ++P.outer;
    P.collisions += static_cast<int>(coll.size());
    // Barrier reflection targets (pipeline.cpp:251-292).
    std::vector<double> blhs(n3 / 3, 0.0), brhs(n3, 0.0);
    for (const Collision& c : coll)
      for (int q = 0; q < 4; ++q) {
        if (c.inv_mass[q] == 0.0f || c.gv[q] >= n3 / 3) continue;
        blhs[c.gv[q]] += kBarrierStiffness;
        for (int d = 0; d < 3; ++d)
          brhs[3 * c.gv[q] + d] +=
              kBarrierStiffness * (curr[3 * c.gv[q] + d] + c.corr[q][d]);
      }
    std::vector<float> next(n3);
    for (int q = 0; q < n3; ++q)
      next[q] = curr[q] + h * vel[q] + h * h * P.acc[q % 3];
    // Outer rhs and (re)factorisation with the barrier diagonal
    // (cloth_solver_utils.cpp:117-163).
    std::vector<double> outer_rhs(n3);
    for (int i : cl) {
      Object& o = P.objs[i];
      std::vector<double> dadd(o.nv, 0.0);
      bool barrier = false;
      for (int v = 0; v < o.nv; ++v) {
        const int g = o.off + v;
        const double m = o.mass[v];
        for (int d = 0; d < 3; ++d) {
          const int q = 3 * g + d;
          outer_rhs[q] = init_rhs[q] + m / (h * h) * curr[q] + m / h * vel[q] +
                         m * P.acc[d] + brhs[q];
        }
        dadd[v] = blhs[g];
        barrier = barrier || blhs[g] != 0.0;
      }
      // The factor holds H, or H plus the last barrier diagonal.
      if (barrier) {
        if (!o.chol.factor(o.H, dadd)) return false;
        o.plain = false;
      } else if (!o.plain) {
        if (!o.chol.factor(o.H, dadd)) return false;
        o.plain = true;
      }
    }
    // Inner loop: projections, global solve (cloth_solver_utils.cpp:205-294,
    // pipeline.cpp:96-121).
    for (int in = 0; in < P.max_inner; ++in) {
      ++P.inner;
      std::vector<float> solution(n3);
      for (int i : cl) {
        Object& o = P.objs[i];
        const int nf = static_cast<int>(o.F.size() / 3);
        std::vector<std::vector<double>> acc(P.threads);
        parallel_for(P.threads, nf, [&](int64_t b, int64_t e, int t) {
          std::vector<double>& rhs = acc[t];
          rhs.assign(3 * static_cast<size_t>(o.nv), 0.0);
          for (int64_t f = b; f < e; ++f) {
            float buf[9];
            for (int a = 0; a < 3; ++a)
              for (int d = 0; d < 3; ++d) buf[3 * a + d] = next[3 * (o.off + o.F[3 * f + a]) + d];
            double D[6];
            for (int r = 0; r < 6; ++r) {
              float s = 0.0f;
              for (int c = 0; c < 9; ++c) s += o.jop[f][9 * r + c] * buf[c];
              D[r] = s;
            }
            double T[6];
            project_stretch(D, 0.9, 1.1, T);
            const float w = P.k * o.area[f];
            for (int c = 0; c < 9; ++c) {
              float s = 0.0f;
              for (int r = 0; r < 6; ++r) s += o.jop[f][9 * r + c] * static_cast<float>(T[r]);
              rhs[3 * o.F[3 * f + c / 3] + c % 3] += w * s;
            }
          }
        });
        std::vector<double> rhs(3 * static_cast<size_t>(o.nv));
        for (int q = 0; q < 3 * o.nv; ++q) rhs[q] = outer_rhs[3 * o.off + q];
        for (const auto& a : acc)
          if (!a.empty())
            for (int q = 0; q < 3 * o.nv; ++q) rhs[q] += a[q];
        std::vector<double> x;
        o.chol.solve3(rhs, x);
        for (int q = 0; q < 3 * o.nv; ++q) solution[3 * o.off + q] = static_cast<float>(x[q]);
      }
      double dn = 0.0;
      for (int q = 0; q < n3; ++q) {
        if (!std::isfinite(solution[q])) return false;
        const double t = solution[q] - next[q];
        dn += t * t;
      }
      next = solution;
      if (std::sqrt(dn) <= 0.05 * scene_scale) break;
    }
    // Project barrier dofs to their targets (pipeline.cpp:294-319).
    if (!coll.empty())
      for (int g = 0; g < n3 / 3; ++g)
        if (blhs[g] != 0.0)
          for (int d = 0; d < 3; ++d)
            next[3 * g + d] = static_cast<float>(brhs[3 * g + d] / blhs[g]);
    // CCD over the segment curr -> next (collision/pipeline.cpp:160-210).
    const float time_start = 1.0f - remaining;
    float toi = remaining;
    {
      const Frame fr = frame(curr, next);
      const std::vector<Candidate> cand = candidates(P, fr);
      P.ccd_candidates += static_cast<int>(cand.size());
      // Scene and collider extents for the numerical filter
      // (collision/pipeline.cpp:166-181).
      double abs_max[3] = {0, 0, 0};
      for (int d = 0; d < 3; ++d)
        abs_max[d] = std::max(std::fabs(bmin[d]), std::fabs(bmax[d]));
      for (int q = 0; q < n3; ++q)
        abs_max[q % 3] = std::max({abs_max[q % 3], static_cast<double>(std::fabs(curr[q])),
                                   static_cast<double>(std::fabs(next[q]))});
      for (const auto& o : P.objs)
        if (!o.dynamic)
          for (int q = 0; q < 3 * o.nv; ++q)
            abs_max[q % 3] = std::max({abs_max[q % 3], static_cast<double>(std::fabs(o.prev[q])),
                                       static_cast<double>(std::fabs(o.curr[q]))});
      double err_vf[3], err_ee[3];
      ccd_filter(abs_max, true, err_vf);
      ccd_filter(abs_max, false, err_ee);
      std::atomic<float> earliest{remaining};
      parallel_for(P.threads, static_cast<int64_t>(cand.size()),
                   [&](int64_t b, int64_t e, int) {
                     for (int64_t i = b; i < e; ++i) {
                       const Candidate& c = cand[i];
                       const bool vf = c.a->type == 2;
                       float p0[4][3], p1[4][3];
                       auto take = [&](int obj, int vtx, int slot) {
                         const Object& o = P.objs[obj];
                         const float* a = fr.t0[obj] + 3 * vtx;
                         const float* bb = fr.t1[obj] + 3 * vtx;
                         for (int d = 0; d < 3; ++d) {
                           if (o.dynamic) {
                             p0[slot][d] = a[d];
                             p1[slot][d] = bb[d];
                           } else {
                             const float delta = bb[d] - a[d];
                             p0[slot][d] = a[d] + time_start * delta;
                             p1[slot][d] = p0[slot][d] + delta;
                           }
                         }
                       };
                       if (vf) {
                         take(c.ob, c.b->v[0], 0);
                         for (int q = 0; q < 3; ++q) take(c.oa, c.a->v[q], 1 + q);
                       } else {
                         take(c.oa, c.a->v[0], 0);
                         take(c.oa, c.a->v[1], 1);
                         take(c.ob, c.b->v[0], 2);
                         take(c.ob, c.b->v[1], 3);
                       }
                       float err[3];
                       for (int d = 0; d < 3; ++d)
                         err[d] = static_cast<float>(vf ? err_vf[d] : err_ee[d]);
                       const float max_t = earliest.load(std::memory_order_relaxed);
                       float t;
                       if (!ccd_query_f32(vf, p0, p1, ms, P.ccd_tol, P.ccd_max_iter, err,
                                          max_t, &t))
                         continue;
                       if (t == 0.0f) continue;
                       float prev = earliest.load(std::memory_order_relaxed);
                       while (t < prev &&
                              !earliest.compare_exchange_weak(prev, t, std::memory_order_relaxed)) {
                       }
                     }
                   });
      toi = earliest.load();
    }
    for (int q = 0; q < n3; ++q) vel[q] = (next[q] - curr[q]) / h;
    if (toi >= remaining) {
      for (int q = 0; q < n3; ++q) curr[q] += remaining * (next[q] - curr[q]);
      break;
    }
    P.min_toi = std::min(P.min_toi, static_cast<double>(time_start + toi));
    const float rb = 0.8f * toi;
    for (int q = 0; q < n3; ++q) curr[q] = rb * (next[q] - curr[q]) + curr[q];
    remaining -= rb;
    coll = find_collisions(1.0f - remaining);
  }
  for (int i : cl) {
    Object& o = P.objs[i];
    for (int q = 0; q < 3 * o.nv; ++q) {
      o.x[q] = curr[3 * o.off + q];
      o.v[q] = vel[3 * o.off + q];
    }
  }
  for (auto& o : P.objs)
    if (!o.dynamic) o.prev = o.curr;
  return true;
}

}  // namespace cpu
}  // namespace oracle

using oracle::cpu::Pipeline;

extern "C" {

/// params: dt, ax, ay, az, k, kb, density, damping, max_outer, max_inner,
/// ms_scale, dcd_scale, ccd_tol, ccd_max_iter.
void* cpu_pipeline_create(int threads, const double* p) {
  auto* P = new Pipeline();
  P->threads = std::max(1, threads);
  P->dt = static_cast<float>(p[0]);
  for (int d = 0; d < 3; ++d) P->acc[d] = static_cast<float>(p[1 + d]);
  P->k = static_cast<float>(p[4]);
  P->kb = static_cast<float>(p[5]);
  P->density = static_cast<float>(p[6]);
  P->damping = static_cast<float>(p[7]);
  P->max_outer = static_cast<int>(p[8]);
  P->max_inner = static_cast<int>(p[9]);
  P->ms_scale = static_cast<float>(p[10]);
  P->dcd_scale = static_cast<float>(p[11]);
  P->ccd_tol = static_cast<float>(p[12]);
  P->ccd_max_iter = static_cast<int>(p[13]);
  return P;
}

void cpu_pipeline_destroy(void* p) { delete static_cast<Pipeline*>(p); }

int cpu_pipeline_add_cloth(void* pp, const float* V, int nv, const int* F, int nf,
                           const int* pins, int npin) {
  Pipeline& P = *static_cast<Pipeline*>(pp);
  oracle::cpu::Object o;
  o.dynamic = true;
  o.nv = nv;
  o.V0.assign(V, V + 3 * nv);
  o.F.assign(F, F + 3 * nf);
  o.pinned.assign(nv, 0);
  for (int i = 0; i < npin; ++i) {
    o.pinned[pins[i]] = 1;
    o.pins.push_back(pins[i]);
    for (int d = 0; d < 3; ++d) o.pin_pos.push_back(V[3 * pins[i] + d]);
  }
  oracle::cpu::prepare_cloth(P, o);
  oracle::cpu::make_colliders(o);
  o.x = o.V0;
  o.v.assign(3 * nv, 0.0f);
  // Cloths before obstacles.
  auto it = std::find_if(P.objs.begin(), P.objs.end(),
                         [](const oracle::cpu::Object& x) { return !x.dynamic; });
  P.objs.insert(it, std::move(o));
  return 0;
}

int cpu_pipeline_add_obstacle(void* pp, const float* V, int nv, const int* F, int nf) {
  Pipeline& P = *static_cast<Pipeline*>(pp);
  oracle::cpu::Object o;
  o.dynamic = false;
  o.nv = nv;
  o.V0.assign(V, V + 3 * nv);
  o.F.assign(F, F + 3 * nf);
  o.pinned.assign(nv, 1);
  oracle::cpu::make_colliders(o);
  o.prev = o.V0;
  o.curr = o.V0;
  P.objs.push_back(std::move(o));
  return 0;
}

/// stats: outer iterations, inner iterations, collisions (summed over outer
/// iterations), CCD candidates, earliest accepted TOI.  Returns 0, or 1 when
/// a factorisation failed or the state became non-finite.
int cpu_pipeline_step(void* pp, double* stats) {
  Pipeline& P = *static_cast<Pipeline*>(pp);
  const bool ok = oracle::cpu::step(P);
  if (stats) {
    stats[0] = P.outer;
    stats[1] = P.inner;
    stats[2] = P.collisions;
    stats[3] = P.ccd_candidates;
    stats[4] = P.min_toi;
  }
  return ok ? 0 : 1;
}

/// Positions and velocities of cloth `idx` (insertion order), 3 per vertex.
void cpu_pipeline_state(void* pp, int idx, float* x, float* v, int set) {
  Pipeline& P = *static_cast<Pipeline*>(pp);
  auto& o = P.objs[idx];
  for (int q = 0; q < 3 * o.nv; ++q) {
    if (set) {
      o.x[q] = x[q];
      o.v[q] = v[q];
    } else {
      x[q] = o.x[q];
      if (v) v[q] = o.v[q];
    }
  }
}

}  // extern "C"
